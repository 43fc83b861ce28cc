"""Device front profile (kg_front_profile, kg_batch_front_profile) against
morphology.radius_profile on the same fields; both are the same FP64
ray march, so they agree to rounding (tolerance 1e-6 cells) and give the
same arm counts."""
import numpy as np
import pytest

from crystalgrowth_amd import morphology

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def kb():
    import crystalgrowth_amd as kb
    if kb.device_count() < 1:
        pytest.fail("no GPU visible")
    return kb


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_profile_matches_host(kb, prec):
    n = 160
    p = kb.default_params(delta=0.04, aniso=6.0, K=2.0, theta0=0.2, noise_a=0.01, boundary=1, seed=2)
    with kb.Solver(n, n, prec, p) as s:
        s.reset_disk(n // 2, n // 2, 5.0)
        for steps in (0, 300, 900):
            s.step(steps)
            phi = s.phase()
            for (cx, cy, k) in ((n // 2, n // 2, 720), (n / 2 + 0.3, n / 2 - 0.7, 97)):
                dev = s.front_profile(cx, cy, k)
                host = morphology.radius_profile(phi, cx, cy, k)
                np.testing.assert_allclose(dev, host, rtol=0, atol=1e-6)
            d = morphology.device_summary(s, n // 2, n // 2)
            h = morphology.summary(phi, n // 2, n // 2)
            assert d["arm_count"] == h["arm_count"]
            assert abs(d["tip_radius"] - h["tip_radius"]) < 1e-6
            assert d["solid_fraction"] == pytest.approx(h["solid_fraction"], abs=1e-12)


def test_batch_profile_matches_host(kb):
    sizes = [(96, 80), (130, 64)]
    params = [kb.default_params(noise_a=0.01, boundary=1, seed=k, instance=k) for k in range(2)]
    with kb.Batch(sizes, params, "f32") as b:
        b.reset_disk(-1, -1, -1, 5.0)
        b.step(200)
        for m, (nx, ny) in enumerate(sizes):
            dev = b.front_profile(m, nx // 2, ny // 2, 360)
            host = morphology.radius_profile(b.phase(m), nx // 2, ny // 2, 360)
            np.testing.assert_allclose(dev, host, rtol=0, atol=1e-6)


def test_profile_rejects_bad_arguments(kb):
    with kb.Solver(32, 32, "f32", kb.default_params()) as s:
        with pytest.raises(kb.KGError):
            s.front_profile(16, 16, 0)
        with pytest.raises(kb.KGError):
            s.front_profile(float("nan"), 16, 8)
