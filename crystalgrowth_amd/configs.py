"""The five benchmark workloads of the Kobayashi solver, defined once.

  c0  300^2 FP64, delta 0.02, j 4, theta0 0.2, noise 0.01, zero-flux, disk r=5, 2000 steps
  c1  4096^2 FP32, j 6, delta 0.04, K 2.0 (else as c0), nine r=5 nuclei, 20000 steps
  c2  512 members, sides drawn from {128, 192, ..., 1024}, per-member delta, K, j,
      theta0, eps_bar; FP32; 5000 steps
  c3  512^2 FP32 adjoint over 1000 steps (checkpoint 50) of L = sum (phi - target)^2,
      target from a run with delta 0.04, K 1.8
  c4  64 members, sides from {128, ..., 512}, hidden (delta, K, theta0) per member,
      500 steps per forward, 50 gradient-descent iterations
Common physics (all but the overrides): dx 0.03, dt 1e-4, tau 3e-4, eps_bar 0.01,
alpha 0.9, gamma 10, Teq 1.
"""
from __future__ import annotations

import numpy as np

from ._lib import Params
from .kobayashi import default_params

C0_PHYSICS = dict(delta=0.02, aniso=4.0, theta0=0.2, noise_a=0.01, boundary=1,
                  dx=0.03, dy=0.03, dt=1e-4, tau=3e-4, eps_bar=0.01, alpha=0.9, gamma=10.0,
                  K=1.6, teq=1.0)


def c0_params(seed: int = 0) -> Params:
    return default_params(**C0_PHYSICS, seed=seed)


def c1_params(seed: int = 0) -> Params:
    return c0_params(seed).replace(aniso=6.0, delta=0.04, K=2.0)


C0 = dict(n=300, steps=2000, precision="f64", disk_r=5.0)
C1 = dict(n=4096, steps=20000, precision="f32", nuclei=9, r=5.0, margin=8, seed=0)


def c2_members(count: int = 512, seed: int = 0):
    """Sizes and kg_params of the ensemble (drawn with numpy's PCG64 from `seed`)."""
    rng = np.random.default_rng(seed)
    sides = rng.choice(np.arange(128, 1025, 64), count)
    sizes, params = [], []
    for m in range(count):
        params.append(c0_params(seed=1000 + m).replace(
            delta=float(rng.uniform(0.01, 0.06)), K=float(rng.uniform(0.8, 2.0)),
            aniso=float(rng.choice([4.0, 5.0, 6.0])), theta0=float(rng.uniform(0, np.pi / 2)),
            eps_bar=float(rng.uniform(0.008, 0.012)), instance=m))
        sizes.append((int(sides[m]), int(sides[m])))
    return sizes, params


C2 = dict(count=512, steps=5000, precision="f32")
C3 = dict(n=512, steps=1000, every=50, precision="f32",
          target=dict(delta=0.04, K=1.8), guess=dict(delta=0.03, K=1.6))


def c3_params(which: str = "guess") -> Params:
    return c0_params(seed=7).replace(**C3[which])


def c4_members(count: int = 64, seed: int = 1):
    """Sizes, hidden (target) parameters and starting guesses of the fit."""
    rng = np.random.default_rng(seed)
    sides = rng.choice(np.arange(128, 513, 64), count)
    sizes, hidden, guess = [], [], []
    for m in range(count):
        base = c0_params(seed=2000 + m).replace(instance=m, aniso=6.0)
        hidden.append(base.replace(delta=float(rng.uniform(0.02, 0.06)),
                                   K=float(rng.uniform(1.2, 2.0)),
                                   theta0=float(rng.uniform(0, np.pi / 3))))
        guess.append(base.replace(delta=0.04, K=1.6, theta0=0.5))
        sizes.append((int(sides[m]), int(sides[m])))
    return sizes, hidden, guess


C4 = dict(count=64, steps=500, iterations=50, every=50, precision="f32")
