"""Host logic of the batched parameter fit (no GPU)."""
import numpy as np

from crystalgrowth_amd import fit
from crystalgrowth_amd.kobayashi import default_params


def test_theta0_error_is_modular_in_the_anisotropy_period():
    h = default_params(aniso=6.0, theta0=0.1, delta=0.04, K=1.6)
    per = 2 * np.pi / 6.0
    same = h.replace(theta0=0.1 + per)           # the same eps(theta) everywhere
    near = h.replace(theta0=0.1 + per - 0.02)    # 0.02 rad the other way round
    e = fit.param_errors([same, near], [h, h])
    assert abs(e["theta0_abs_rad"] - 0.01) < 1e-12
    assert e["delta_rel"] == 0.0 and e["K_rel"] == 0.0


def test_relative_errors():
    h = default_params(delta=0.04, K=1.6)
    p = h.replace(delta=0.05, K=1.2)
    e = fit.param_errors([p], [h])
    assert abs(e["delta_rel"] - 0.25) < 1e-12 and abs(e["K_rel"] - 0.25) < 1e-12


def test_c4_members_are_reproducible():
    from crystalgrowth_amd import configs
    a = configs.c4_members(8)
    b = configs.c4_members(8)
    assert a[0] == b[0]
    assert all(128 <= n <= 512 and n % 64 == 0 for n, _ in a[0])
    for h, g in zip(a[1], a[2]):
        assert 0.02 <= h.delta <= 0.06 and 1.2 <= h.K <= 2.0 and 0 <= h.theta0 <= np.pi / 3
        assert g.instance == h.instance


class QuadraticBatch:
    """Stands in for kg_batch in the optimiser: member m's loss is a
    quadratic bowl around its hidden parameters (no GPU)."""

    def __init__(self, hidden, weights):
        self.hidden, self.w = hidden, weights
        self.cur = [None] * len(hidden)
        self.calls = 0

    def set_params(self, m, p):
        self.cur[m] = p

    def reset_disk(self, *args):
        pass

    def adjoint(self, steps, targets, every):
        self.calls += 1
        out = []
        for p, h, w in zip(self.cur, self.hidden, self.w):
            r = np.array([getattr(p, n) - getattr(h, n) for n in fit.NAMES])
            out.append({"loss": float((w * r * r).sum()),
                        **{"d_" + n: float(2 * w[k] * r[k]) for k, n in enumerate(fit.NAMES)}})
        return out, None


def test_bfgs_reaches_the_minimum_of_ill_scaled_bowls():
    """Bowls whose curvatures differ by 1e6 across the parameters (the
    loss is far flatter in delta than in K or theta0): the batched
    quasi-Newton fit lands every member on its hidden parameters within the
    configuration's 50 iterations, one batched evaluation each."""
    from crystalgrowth_amd import configs
    _, hidden, guess = configs.c4_members(8)
    w = [np.array([3e2, 1e5, 3e4]) * (1 + 0.3 * k) for k in range(8)]
    b = QuadraticBatch(hidden, w)
    cur, losses = fit._fit_bfgs(b, guess, None, 500, 50, 50, 5.0, None)
    assert b.calls == 50 and len(losses) == 50 and losses[-1] < 1e-6 * losses[0]
    e = fit.param_errors(cur, hidden)
    assert e["delta_rel"] < 1e-3 and e["K_rel"] < 1e-4 and e["theta0_abs_rad"] < 1e-4
