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
