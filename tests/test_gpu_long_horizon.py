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


def test_c1_physics_morphology_tip_velocity(kb, O):
    """The 4096^2 configuration's physics (zero-flux, j = 6, delta = 0.04,
    K = 2.0, theta0 = 0.2, noise 0.01) on 768^2 from one r = 5 disk at the
    centre: FP32 GPU against the FP32 oracle after 3000 and 4000 steps.
    Tolerances: solid fraction within 2 %, tip radius within 3 %, tip velocity
    (3000 -> 4000) within 5 %, the same arm count (6)."""
    n, s1, s2 = 768, 3000, 4000
    p = O.default_params(delta=0.04, aniso=6.0, K=2.0, theta0=0.2, noise_a=0.01, boundary=1,
                         seed=0)
    a = O.run_parallel(O.init_disk(n, n, n // 2, n // 2, 5.0, np.float32), p, s1)
    b = O.run_parallel(a, p, s2 - s1, step0=s1)
    with kb.Solver(n, n, "f32", kb.Params(**p.as_dict())) as g:
        g.reset_disk(n // 2, n // 2, 5.0)
        g.step(s1)
        ga = g.phase()
        g.step(s2 - s1)
        gb = g.phase()
        dev = M.device_summary(g, n // 2, n // 2)
    c = n // 2
    for got, ref in ((ga, a.phi), (gb, b.phi)):
        x, y = M.summary(got, c, c), M.summary(ref, c, c)
        assert abs(x["solid_fraction"] - y["solid_fraction"]) <= 0.02 * y["solid_fraction"]
        assert abs(x["tip_radius"] - y["tip_radius"]) <= 0.03 * y["tip_radius"]
        assert x["arm_count"] == y["arm_count"] == 6
    span = (s2 - s1) * p.dt
    vg = M.tip_velocity(ga, gb, c, c, span, p.dx)
    vo = M.tip_velocity(a.phi, b.phi, c, c, span, p.dx)
    assert vo > 0 and abs(vg - vo) <= 0.05 * vo, (vg, vo)
    # the device's morphology of the final state agrees with the host's
    assert dev["arm_count"] == 6
    assert dev["tip_radius"] == pytest.approx(M.tip_radius(gb, c, c), abs=0.3)


@pytest.mark.parametrize("precision,tol", [("f32", 2e-3), ("f64", 1e-11)])
def test_headline_invariant_at_full_size(kb, precision, tol):
    """The bench's 4096 x 4096 workload over the first 10000 steps of its
    horizon: sum(T - K phi) is conserved by the model (the 9-point stencil's
    weights telescope and the noise enters phi and T together), so it may
    move only by accumulated rounding (measured: 4.1e-4 of its value in
    FP32, 3.4e-14 in FP64), while a dropped or mis-signed term of either update, or a K off
    by 1 %, moves it by more than its whole value (sum(phi) grows from 729
    to about 5.7e5 over these steps).  The fields stay finite and the
    crystals grow."""
    p = kb.default_params(delta=0.04, aniso=6.0, K=2.0, theta0=0.2, noise_a=0.01,
                          boundary=1, seed=0)
    with kb.Solver(4096, 4096, precision, p) as s:
        s.reset_seeds(9, 5.0, 8, 0)
        st0 = s.stats()
        s.step(10000)
        st = s.stats()
    inv0, inv = st0["sum_T_minus_K_phi"], st["sum_T_minus_K_phi"]
    print("sum(T - K phi):", inv0, "->", inv, "relative drift", abs(inv - inv0) / abs(inv0))
    assert st["nonfinite"] == 0
    assert st["solid_fraction"] > 5 * st0["solid_fraction"]
    assert abs(inv - inv0) <= tol * abs(inv0)
