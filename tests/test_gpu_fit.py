"""Batched parameter fit on the GPU: a few Adam iterations of kg_batch_adjoint
lower the loss of every member set (small members, short horizon)."""
import pytest

pytestmark = pytest.mark.gpu


def test_fit_lowers_the_loss():
    import crystalgrowth_amd as kb
    from crystalgrowth_amd import configs, fit
    if kb.device_count() < 1:
        pytest.fail("no GPU visible")
    sizes, hidden, guess = configs.c4_members(6, seed=3)
    sizes = [(64 + 16 * (k % 3), 64) for k in range(6)]
    res = fit.fit(sizes, hidden, guess, steps=120, iterations=6, every=40)
    assert res["loss_last"] < 0.7 * res["loss_first"], res
    assert res["param_error_end"]["K_rel"] < res["param_error_start"]["K_rel"], res
