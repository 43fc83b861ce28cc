"""Long horizons (needs a GPU): dendritic growth amplifies ulp-level
differences, so the GPU run and the CPU oracle are compared on morphology.

Tolerances: solid fraction within 2 % relative, tip radius within 3 %,
identical arm count; plus the FP64 kernel against the FP64 oracle per cell
(1e-9) on a disk-seeded run of the same length, where no symmetry point makes
the angle update ill-conditioned.
"""
import numpy as np
import pytest

from crystalgrowth_amd import morphology as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def kb():
    import crystalgrowth_amd as kb
    if kb.device_count() < 1:
        pytest.fail("no GPU visible")
    return kb


def test_reference_default_run_morphology(kb, O):
    """src/main.cpp's run: 250^2, dt = 1e-4, plus nucleus, periodic, 200 frames
    of 10 steps, in the reference's float arithmetic."""
    n, steps = 250, 2000
    ref = O.run(O.init_plus(n, n, np.float32), O.default_params(), steps)
    k = kb.Kobayashi(n, n, 0.0001)
    for _ in range(steps // 10):
        k.iUpdate()
    phi = k.phase()
    k.close()
    a, b = M.summary(phi, n // 2, n // 2), M.summary(ref.phi, n // 2, n // 2)
    assert abs(a["solid_fraction"] - b["solid_fraction"]) <= 0.02 * b["solid_fraction"]
    assert abs(a["tip_radius"] - b["tip_radius"]) <= 0.03 * b["tip_radius"]
    assert a["arm_count"] == b["arm_count"] == 6


def test_fp64_disk_run_per_cell_2000_steps(kb, O):
    n, steps = 128, 2000
    p = O.default_params(delta=0.04, aniso=6.0, K=2.0, theta0=0.2, noise_a=0.01, boundary=1, seed=9)
    s = O.init_disk(n, n, 60, 66, 5.0)
    ref = O.run(s, p, steps)
    with kb.Solver(n, n, "f64", kb.Params(**p.as_dict())) as g:
        g.set_fields(s.phi, s.T, s.theta)
        g.step(steps)
        phi, T, _ = g.fields()
    assert np.abs(phi - ref.phi).max() <= 1e-9
    assert np.abs(T - ref.T).max() <= 1e-9
