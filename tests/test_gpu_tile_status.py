"""Tile status of the fused kernel (needs a GPU).

The default fused step keeps one status word per tile and buffer (phi all +0,
phi all 1, T all +0, tile took the full path) and uses it to skip loads and
stores whose values are known, to skip tiles whose 3 x 3 neighbourhood is
all zero, and to stage theta early on active tiles.  None of this may change
a bit: every comparison here is bitwise against KG_KERNEL_FUSED_DENSE (the
same launch without status).  Grids are cut into partial tiles on purpose,
and the runs start from nuclei in a zero field so that all tile kinds occur
(quiescent, phi known, T known, flat, active)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def kb():
    import crystalgrowth_amd as kb
    if kb.device_count() < 1:
        pytest.fail("no GPU visible")
    return kb


def run_pair(kb, nx, ny, prec, p, setup, marks, stats=None):
    """Advance a status solver and a dense solver side by side; fields at each
    mark (and the status solver's skip statistics appended to stats)."""
    out = []
    with kb.Solver(nx, ny, prec, p) as a, kb.Solver(nx, ny, prec, p) as b:
        b.set_kernel("fused_dense")
        setup(a)
        setup(b)
        done = 0
        for m in marks:
            a.step(m - done)
            b.step(m - done)
            done = m
            out.append((m, a.fields(), b.fields()))
        if stats is not None:
            stats.append(a.skip_stats())
    return out


def assert_bitwise(pairs):
    for m, fa, fb in pairs:
        for name, x, y in zip(("phi", "T", "theta"), fa, fb):
            assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), (m, name)


@pytest.mark.parametrize("boundary", [0, 1])
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_status_path_matches_dense_bitwise(kb, boundary, prec):
    nx, ny = 460, 250          # 8 x 9 tiles, the last row and column partial
    p = kb.default_params(delta=0.04, aniso=6.0, K=2.0, theta0=0.2, noise_a=0.01,
                          boundary=boundary, seed=3)
    pairs = run_pair(kb, nx, ny, prec, p, lambda s: s.reset_seeds(3, 5.0, 8, 11),
                     [1, 2, 3, 40, 400])
    assert_bitwise(pairs)


def test_periodic_grid_with_thin_edge_tiles(kb):
    # width % 64 == 1 and height % 30 == 1: no status is kept (dense launch)
    nx, ny = 129, 61
    p = kb.default_params(noise_a=0.01, boundary=0, seed=5)
    pairs = run_pair(kb, nx, ny, "f32", p, lambda s: s.reset_disk(nx - 3, ny - 2, 6.0), [1, 60])
    assert_bitwise(pairs)


def test_invalidation_after_external_writes(kb):
    """set_fields, a dense step and a kernel switch in between: the status must
    be rebuilt, not trusted."""
    nx, ny = 300, 200
    p = kb.default_params(noise_a=0.01, boundary=1, seed=9)
    rng = np.random.default_rng(2)
    with kb.Solver(nx, ny, "f32", p) as a, kb.Solver(nx, ny, "f32", p) as b:
        b.set_kernel("fused_dense")
        for s in (a, b):
            s.reset_disk(60, 50, 5.0)
            s.step(30)
        # new fields: zero almost everywhere, a blob where the old state was zero
        phi = np.zeros((ny, nx), np.float32)
        phi[150:160, 200:230] = 1.0
        T = np.zeros((ny, nx), np.float32)
        T[20:40, 250:290] = rng.random((20, 40)).astype(np.float32) * 0.3
        for s in (a, b):
            s.set_fields(phi, T)
            s.step(25)
        assert_bitwise([(55, a.fields(), b.fields())])
        a.set_kernel("two_pass")
        a.step(1)
        a.set_kernel("fused")
        b.set_kernel("two_pass")
        b.step(1)
        b.set_kernel("fused_dense")
        for s in (a, b):
            s.step(20)
        assert_bitwise([(76, a.fields(), b.fields())])


def test_batch_members_match_dense_solver(kb):
    sizes = [(200, 130), (64, 64), (130, 61), (96, 250)]
    params = [kb.default_params(delta=0.04, aniso=6.0, noise_a=0.01, boundary=k % 2, seed=k,
                                instance=k) for k in range(len(sizes))]
    with kb.Batch(sizes, params, "f32") as bt:
        bt.reset_disk(-1, -1, -1, 5.0)
        bt.step(150)
        got = [bt.fields(m) for m in range(len(sizes))]
    for (nx, ny), p, g in zip(sizes, params, got):
        with kb.Solver(nx, ny, "f32", p) as s:
            s.set_kernel("fused_dense")
            s.reset_disk(nx // 2, ny // 2, 5.0)
            s.step(150)
            assert_bitwise([(150, g, s.fields())])


@pytest.mark.parametrize("boundary", [0, 1])
def test_scalar_staging_and_tiny_grids(kb, boundary):
    """Widths that are not a multiple of 4 stage cell by cell; 3 x 3 and
    5 x 4 grids are one tile whose neighbours are itself."""
    p = kb.default_params(delta=0.04, aniso=6.0, noise_a=0.01, boundary=boundary, seed=8)
    for nx, ny, setup in ((301, 77, lambda s: s.reset_seeds(2, 4.0, 6, 5)),
                          (3, 3, lambda s: s.reset()),
                          (5, 4, lambda s: s.reset())):
        pairs = run_pair(kb, nx, ny, "f32", p, setup, [1, 3, 60])
        assert_bitwise(pairs)


@pytest.mark.parametrize("boundary,prec", [(1, "f32"), (0, "f32"), (1, "f64")])
def test_chunk_skipping_matches_dense_bitwise(kb, boundary, prec):
    """Far-field chunks are left alone by the status launch (their 2-tile
    neighbourhood is all +0) while one nucleus grows into them.  Periodic:
    the nucleus sits at a corner, so its fields wrap to all four edges.
    Steps are taken in uneven calls, so the host's request periods fall
    across call boundaries."""
    nx, ny = 2048, 1920        # 32 x 64 tiles, 16 x 16 chunks of 2 x 4 tiles
    p = kb.default_params(delta=0.04, aniso=6.0, K=2.0, theta0=0.2, noise_a=0.01,
                          boundary=boundary, seed=4)
    cx, cy = (3, 2) if boundary == 0 else (150, 140)
    stats = []
    pairs = run_pair(kb, nx, ny, prec, p, lambda s: s.reset_disk(cx, cy, 6.0),
                     [1, 6, 13, 31, 90, 301, 700], stats)
    assert_bitwise(pairs)
    # most launches after the first answer cover only the live chunks
    assert stats[0]["listed_launches"] >= 600 and stats[0]["stale_answers"] == 0, stats


def test_small_grid_launches_every_chunk(kb):
    """Below 256 chunks the handle never waits for liveness answers."""
    with kb.Solver(1024, 960, "f32", kb.default_params(noise_a=0.01, seed=2)) as s:
        s.reset_disk(100, 100, 5.0)
        s.step(100)
        s.synchronize()
        assert s.skip_stats() == {"listed_launches": 0, "stale_answers": 0}


def test_chunk_skipping_after_new_fields(kb):
    """A blob written far from the old activity while skip answers about the
    old fields are still in flight: the answers must be dropped."""
    nx, ny = 2048, 1920
    p = kb.default_params(noise_a=0.01, boundary=1, seed=6)
    with kb.Solver(nx, ny, "f32", p) as a, kb.Solver(nx, ny, "f32", p) as b:
        b.set_kernel("fused_dense")
        for s in (a, b):
            s.reset_disk(100, 100, 5.0)
            s.step(41)
        phi = np.zeros((ny, nx), np.float32)
        phi[1700:1712, 1800:1815] = 1.0
        T = np.zeros((ny, nx), np.float32)
        T[1500:1520, 20:60] = 0.2
        for s in (a, b):
            s.set_fields(phi, T)
            s.step(3)
            s.step(30)
        assert_bitwise([(74, a.fields(), b.fields())])


def test_headline_trajectory_matches_dense_bitwise(kb):
    """The 4096 x 4096 workload of the bench line (zero-flux, noise 0.01,
    nine r = 5 seeds) over its 20200 steps: every skip (tile status, chunk
    lists ordered heaviest first, liveness requests carried by the step
    launches) leaves the final fields bitwise those of the dense kernel."""
    p = kb.default_params(delta=0.04, aniso=6.0, K=2.0, theta0=0.2, noise_a=0.01,
                          boundary=1, seed=0)
    stats = []
    pairs = run_pair(kb, 4096, 4096, "f32", p, lambda s: s.reset_seeds(9, 5.0, 8, 0),
                     [200, 20200], stats)
    assert_bitwise(pairs)
    assert stats[0]["listed_launches"] >= 20000 and stats[0]["stale_answers"] == 0, stats


@pytest.mark.parametrize("boundary", [0, 1])
@pytest.mark.parametrize("nx,ny,prec", [(2050, 1930, "f32"), (2051, 1931, "f32"),
                                        (2050, 1930, "f64")])
def test_chunk_skipping_on_ragged_chunk_grid(kb, boundary, nx, ny, prec):
    """2050 x 1930 cells: 33 x 65 tiles, the last tile column 2 cells wide and
    the last tile row 10 rows high, so the last chunk column and row are
    partial (17 x 17 chunks of 2 x 4 tiles).  A nucleus next to the far
    corner grows into the partial chunks; periodic fields wrap across it.
    2051 x 1931 (3 and 11 cells) also takes the cell-by-cell staging of
    widths that are not a multiple of 4."""
    p = kb.default_params(delta=0.04, aniso=6.0, K=2.0, theta0=0.2, noise_a=0.01,
                          boundary=boundary, seed=8)
    stats = []
    pairs = run_pair(kb, nx, ny, prec, p, lambda s: s.reset_disk(nx - 20, ny - 15, 6.0),
                     [1, 8, 29, 150, 400], stats)
    assert_bitwise(pairs)
    assert stats[0]["listed_launches"] >= 350 and stats[0]["stale_answers"] == 0, stats


def test_chunk_skipping_across_step_count_changes(kb):
    """The step count (the noise counter) is set between calls, off the
    liveness period, while an answer is in flight and a list is in use; the
    status solver must still match the dense one bitwise."""
    nx, ny = 2048, 1920
    p = kb.default_params(delta=0.04, aniso=6.0, K=2.0, theta0=0.2, noise_a=0.01,
                          boundary=0, seed=9)
    with kb.Solver(nx, ny, "f32", p) as a, kb.Solver(nx, ny, "f32", p) as b:
        b.set_kernel("fused_dense")
        for s in (a, b):
            s.reset_disk(1000, 900, 8.0)
        for count, n in ((None, 60), (1003, 9), (5, 12), (77777, 40)):
            for s in (a, b):
                if count is not None:
                    s.step_count = count
                s.step(n)
        assert_bitwise([(121, a.fields(), b.fields())])
        assert a.skip_stats()["listed_launches"] > 0


def test_chunk_lists_longer_than_kernel_arguments(kb):
    """More than 1020 live chunks (the kernel arguments' list capacity): the
    list goes to the launch through device memory, and the run must still
    match the dense kernel bitwise.  6144 x 4096 cells are 48 x 35 chunks
    (1680); 800 nuclei spread over the grid keep about 1280 of them live
    from the start."""
    nx, ny = 6144, 4096
    p = kb.default_params(delta=0.04, aniso=6.0, K=2.0, theta0=0.2, noise_a=0.01,
                          boundary=1, seed=9)
    with kb.Solver(nx, ny, "f32", p) as s:
        s.reset_seeds(800, 5.0, 8, 21)
        phi0 = s.phase()
    # live chunks of the initial state: a chunk (2 x 4 tiles of 64 x 30) is
    # live when a tile within one tile column / row of it holds phi != 0
    tiles_y, tiles_x = -(-ny // 30), -(-nx // 64)
    busy = np.zeros((tiles_y, tiles_x), bool)
    ys, xs = np.nonzero(phi0)
    busy[ys // 30, xs // 64] = True
    live = 0
    for cy in range(-(-tiles_y // 4)):
        for cx in range(tiles_x // 2):
            y0, x0 = max(cy * 4 - 1, 0), max(cx * 2 - 1, 0)
            live += bool(busy[y0:cy * 4 + 5, x0:cx * 2 + 3].any())
    assert 1020 < live < 1680, live
    stats = []
    pairs = run_pair(kb, nx, ny, "f32", p, lambda s: s.reset_seeds(800, 5.0, 8, 21),
                     [9, 50], stats)
    assert_bitwise(pairs)
    assert stats[0]["listed_launches"] >= 30 and stats[0]["stale_answers"] == 0, stats
