"""kg_colormap_rgba8_noalloc through the C ABI (needs a GPU): the colours staged
in the library's static device buffer equal kg_colormap_rgba8's bytes, on two
calls in a row whose second grid is smaller than the first, and grids beyond
the stated maximum are refused with KG_ERR_SIZE."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def kb():
    import crystalgrowth_amd as kb
    if kb.device_count() < 1:
        pytest.fail("no GPU visible")
    return kb


def grown(kb, nx, ny, seed):
    s = kb.Solver(nx, ny, "f32", kb.default_params(delta=0.04, noise_a=0.01, seed=seed))
    s.reset_seeds(4, 6.0, 10, seed)
    s.step(60)
    return s


def test_noalloc_matches_allocating_ramp_on_consecutive_calls(kb):
    # the first call fills most of the buffer, the second covers less of it
    for nx, ny, order in ((2048, 1920, 0), (640, 640, 1), (300, 200, 0)):
        s = grown(kb, nx, ny, nx)
        try:
            want = s.colormap_rgba8(order)
            got = s.colormap_rgba8_noalloc(order)
            assert np.array_equal(got, want), (nx, ny, order)
            assert np.unique(want[:, 0]).size > 1      # the nuclei show
        finally:
            s.close()


def test_noalloc_rejects_grids_beyond_its_maximum(kb):
    from crystalgrowth_amd import _lib
    cap = 1 << 22
    with kb.Solver(2049, 2048, "f32") as s:         # 4196352 cells > 4194304
        with pytest.raises(_lib.KGError) as e:
            s.colormap_rgba8_noalloc(0)
        assert e.value.code == _lib.KG_ERR_SIZE and str(cap) in str(e.value)
    with kb.Solver(2048, 2048, "f32") as s:         # exactly the maximum
        assert np.array_equal(s.colormap_rgba8_noalloc(1), s.colormap_rgba8(1))
