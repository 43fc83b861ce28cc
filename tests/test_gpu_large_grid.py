"""A grid whose FP32 byte offsets pass 2^31 (needs a GPU).

24576 x 24000 cells (590 M; the last row starts 2.36 GB into each field;
38400 chunks, so the launches use chunk lists).  A disk next to the far,
zero-flux corner must evolve bitwise as the same disk on a 256 x 256 grid
whose right and bottom edges are that corner's: the step of a cell depends
only on its neighbourhood and the boundary it sees, and without noise not on
its absolute position.  A 32-bit slip in any address (field, theta, status,
chunk list) would break the match or leave the far corner unstepped."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def kb():
    import crystalgrowth_amd as kb
    if kb.device_count() < 1:
        pytest.fail("no GPU visible")
    return kb


def test_far_corner_of_a_large_grid_matches_small_grid(kb):
    NX, NY, S, STEPS = 24576, 24000, 256, 40
    assert (NY - 1) * NX * 4 > 2 ** 31
    p = kb.default_params(delta=0.04, aniso=6.0, K=2.0, theta0=0.2, noise_a=0.0,
                          boundary=1, seed=3)
    cx, cy, r = S - 40, S - 50, 12.0       # 40 / 50 cells from the shared edges
    with kb.Solver(S, S, "f32", p) as s:
        s.reset_disk(cx, cy, r)
        s.step(STEPS)
        small = s.fields()
    with kb.Solver(NX, NY, "f32", p) as b:
        b.reset_disk(NX - S + cx, NY - S + cy, r)
        b.step(STEPS)
        b.synchronize()
        stats = b.skip_stats()
        big = [f[NY - S:, NX - S:].copy() for f in b.fields()]
    assert float(small[0].max()) == 1.0 and np.count_nonzero(small[1]) > 1000
    for name, x, y in zip(("phi", "T", "theta"), big, small):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), name
    assert stats["listed_launches"] >= STEPS - 14, stats
