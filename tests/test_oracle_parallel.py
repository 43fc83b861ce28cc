"""The OpenMP build of the oracle (all-cores CPU timing) is the same
computation: bitwise equal to the serial build, both precisions and both
boundaries, noise on."""
import numpy as np
import pytest


@pytest.mark.parametrize("boundary", [0, 1])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_parallel_oracle_is_bitwise_serial(O, boundary, dtype):
    p = O.default_params(delta=0.04, aniso=6.0, K=2.0, theta0=0.2, noise_a=0.01,
                         boundary=boundary, seed=4)
    s, _ = O.init_seeds(96, 70, 3, 4.0, 6, 2, dtype=dtype)
    a = O.run(s, p, 25, step0=3)
    b = O.run_parallel(s, p, 25, step0=3)
    for x, y in ((a.phi, b.phi), (a.T, b.T), (a.theta, b.theta)):
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8))
    assert O.threads_parallel() >= 1
