"""Batched ensemble (kg_batch_*) on the GPU: every member evolves bit for bit
like a plain handle with the same size, precision, parameters and state, in
per-step-launch and in hipGraph-replay mode."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def kb():
    import crystalgrowth_amd as kb
    if kb.device_count() < 1:
        pytest.fail("no GPU visible")
    return kb


def members(kb):
    sizes = [(128, 128), (37, 23), (200, 96), (64, 64), (130, 70), (5, 4)]
    params = [
        kb.default_params(delta=0.04, aniso=6.0, K=2.0, theta0=0.2, noise_a=0.01, boundary=1, seed=1, instance=0),
        kb.default_params(delta=0.02, aniso=4.0, noise_a=0.02, seed=2, instance=1),
        kb.default_params(delta=0.05, boundary=1, seed=3, instance=2),
        kb.default_params(aniso=5.0, theta0=0.7, noise_a=0.01, seed=4, instance=3),
        kb.default_params(delta=0.06, K=0.9, eps_bar=0.012, boundary=0, noise_a=0.01, seed=5, instance=4),
        kb.default_params(boundary=1),
    ]
    return sizes, params


def initial(sizes, prec):
    dt = np.float32 if prec == "f32" else np.float64
    out = []
    rng = np.random.default_rng(7)
    for k, (nx, ny) in enumerate(sizes):
        if k in (2, 5):
            phi = rng.random((ny, nx))
            T = rng.random((ny, nx)) * 0.5
            th = rng.uniform(0, 6, (ny, nx))
        else:
            yy, xx = np.mgrid[0:ny, 0:nx]
            phi = (((xx - nx // 2) ** 2 + (yy - ny // 2) ** 2) <= 25).astype(float)
            T = np.zeros((ny, nx))
            th = np.zeros((ny, nx))
        out.append(tuple(a.astype(dt) for a in (phi, T, th)))
    return out


def plain(kb, nx, ny, prec, p, state, nsteps):
    with kb.Solver(nx, ny, prec, p) as s:
        s.set_fields(*state)
        s.step(nsteps)
        return s.fields()


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("mode,g", [("launch", 0), ("graph", 4)])
def test_members_match_plain_solver_bitwise(kb, prec, mode, g):
    sizes, params = members(kb)
    init = initial(sizes, prec)
    with kb.Batch(sizes, params, prec) as b:
        if mode == "graph":
            b.set_mode("graph", g)
        for m, st in enumerate(init):
            b.set_fields(m, *st)
        b.step(11)  # 2 replays of 4 + 3 single launches in graph mode
        assert b.step_count == 11
        for m, (nx, ny) in enumerate(sizes):
            ref = plain(kb, nx, ny, prec, params[m], init[m], 11)
            got = b.fields(m)
            for x, y in zip(got, ref):
                assert np.array_equal(x, y), (m, nx, ny)


def test_graph_replays_each_match_eager(kb):
    """Replays with eager readbacks and synchronizations in between: every
    replay advances the noise counter and matches the plain solver."""
    sizes, params = members(kb)
    init = initial(sizes, "f32")
    solvers = []
    for m, (nx, ny) in enumerate(sizes):
        s = kb.Solver(nx, ny, "f32", params[m])
        s.set_fields(*init[m])
        solvers.append(s)
    with kb.Batch(sizes, params, "f32") as b:
        b.set_mode("graph", 6)
        for m, st in enumerate(init):
            b.set_fields(m, *st)
        for rep in range(4):
            b.step(6)
            b.synchronize()
            for m, s in enumerate(solvers):
                s.step(6)
                assert np.array_equal(b.phase(m), s.phase()), (rep, m)
                assert np.array_equal(b.temperature(m), s.temperature()), (rep, m)
            st = b.stats(0)           # eager work between replays
            assert st["nonfinite"] == 0
        # an odd remainder, then graph replays again from odd parity
        b.step(1)
        b.step(6)
        for m, s in enumerate(solvers):
            s.step(7)
            assert np.array_equal(b.phase(m), s.phase()), m
    for s in solvers:
        s.close()


def test_autotune_restores_state(kb):
    sizes, params = members(kb)
    init = initial(sizes, "f32")
    with kb.Batch(sizes, params, "f32") as b:
        for m, st in enumerate(init):
            b.set_fields(m, *st)
        b.step(3)
        before = [b.fields(m) for m in range(len(sizes))]
        mode, g = b.autotune(8)
        assert mode in ("launch", "graph", "graph_lists") and (mode == "launch" or g >= 2)
        assert b.chunk_shape in kb.Batch.CHUNK_SHAPES
        assert b.step_count == 3
        for m in range(len(sizes)):
            for x, y in zip(b.fields(m), before[m]):
                assert np.array_equal(x, y)
        b.step(5)
        for m, (nx, ny) in enumerate(sizes):
            ref = plain(kb, nx, ny, "f32", params[m], init[m], 8)
            assert np.array_equal(b.phase(m), ref[0]), m


def test_resets_and_reduce(kb):
    sizes, params = members(kb)
    with kb.Batch(sizes, params, "f32") as b:
        b.reset_disk(-1, -1, -1, 5.0)
        assert b.step_count == 0
        b.step(20)
        for m, (nx, ny) in enumerate(sizes):
            with kb.Solver(nx, ny, "f32", params[m]) as s:
                s.reset_disk(nx // 2, ny // 2, 5.0)
                s.step(20)
                assert np.array_equal(b.phase(m), s.phase())
                assert b.stats(m) == s.stats()
        b.reset_seeds(3, 2.0, 1, 99, member=0)  # one member: the counter continues
        assert b.step_count == 20
        with kb.Solver(*sizes[0], "f32", params[0]) as s:
            s.reset_seeds(3, 2.0, 1, 99)
            assert np.array_equal(b.phase(0), s.phase())


def test_member_with_wide_stored_angles(kb):
    """A member given angles up to 1e5 (outside the reduced sin/cos range)
    takes the checked path and still matches a plain handle bitwise."""
    nx, ny = 70, 40
    p = kb.default_params(aniso=6.0, theta0=0.3, boundary=1, noise_a=0.01, seed=4)
    rng = np.random.default_rng(5)
    phi = rng.random((ny, nx)).astype(np.float32)
    phi[:, nx // 2:] = 0.3
    T = (rng.random((ny, nx)) * 0.5).astype(np.float32)
    th = rng.uniform(-1e5, 1e5, (ny, nx)).astype(np.float32)
    with kb.Batch([(64, 64), (nx, ny)], [kb.default_params(), p], "f32") as b:
        b.set_fields(1, phi, T, th)
        b.step(20)
        got = b.fields(1)
    ref = plain(kb, nx, ny, "f32", p, (phi, T, th), 20)
    for x, y in zip(got, ref):
        assert np.array_equal(x, y)


def list_batch(kb):
    sizes = [(1024, 960), (770, 610), (1030, 901), (640, 480), (900, 905)]
    params = [
        kb.default_params(delta=0.04, aniso=6.0, K=2.0, theta0=0.2, noise_a=0.01, boundary=1, seed=1, instance=0),
        kb.default_params(delta=0.02, aniso=4.0, noise_a=0.02, boundary=0, seed=2, instance=1),
        kb.default_params(delta=0.05, noise_a=0.01, boundary=0, seed=3, instance=2),
        kb.default_params(aniso=5.0, theta0=0.7, noise_a=0.01, boundary=1, seed=4, instance=3),
        kb.default_params(delta=0.06, K=0.9, noise_a=0.01, boundary=0, seed=5, instance=4),
    ]
    spots = [(100, 120), (766, 3), (5, 895), (320, 240), (899, 450)]
    return sizes, params, spots


def check_against_dense(kb, prec, sizes, params, spots, got, marks):
    for m, ((nx, ny), p, (cx, cy)) in enumerate(zip(sizes, params, spots)):
        with kb.Solver(nx, ny, prec, p) as s:
            s.set_kernel("fused_dense")
            s.reset_disk(cx, cy, 6.0)
            done = 0
            for mk in marks:
                s.step(mk - done)
                done = mk
                for name, x, y in zip(("phi", "T", "theta"), got[mk][m], s.fields()):
                    assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), (m, mk, name)


@pytest.mark.parametrize("prec,shape", [("f32", (2, 2)), ("f64", (2, 2)), ("f32", (2, 4)),
                                        ("f32", (4, 2)), ("f32", (4, 1)), ("f32", (8, 1))])
def test_chunk_lists_match_plain_solver_bitwise(kb, prec, shape):
    """Launch mode over a batch of 256+ chunks takes chunk lists (members'
    far fields left out, heaviest chunks first): every member still evolves
    bit for bit like a dense plain handle, in every chunk shape (the fewest
    chunks, 262 at 2 x 4, still take lists).  Members differ in size,
    boundary and partial edge tiles; nuclei sit off centre so fronts cross
    members' chunk edges and, on periodic members, wrap."""
    sizes = [(1024, 960), (770, 610), (1030, 901), (640, 480), (900, 905)]
    params = [
        kb.default_params(delta=0.04, aniso=6.0, K=2.0, theta0=0.2, noise_a=0.01, boundary=1, seed=1, instance=0),
        kb.default_params(delta=0.02, aniso=4.0, noise_a=0.02, boundary=0, seed=2, instance=1),
        kb.default_params(delta=0.05, noise_a=0.01, boundary=0, seed=3, instance=2),
        kb.default_params(aniso=5.0, theta0=0.7, noise_a=0.01, boundary=1, seed=4, instance=3),
        kb.default_params(delta=0.06, K=0.9, noise_a=0.01, boundary=0, seed=5, instance=4),
    ]
    spots = [(100, 120), (766, 3), (5, 895), (320, 240), (899, 450)]
    marks = [1, 9, 40, 151, 400]
    with kb.Batch(sizes, params, prec) as b:
        assert b.chunk_shape == (2, 2)
        b.set_chunk_shape(*shape)
        assert b.chunk_shape == shape
        b.set_mode("launch")
        for m, (cx, cy) in enumerate(spots):
            b.reset_disk(m, cx, cy, 6.0)
        got, done = {}, 0
        for mk in marks:
            b.step(mk - done)
            done = mk
            got[mk] = [b.fields(m) for m in range(len(sizes))]
        stats = b.skip_stats()
    for m, ((nx, ny), p, (cx, cy)) in enumerate(zip(sizes, params, spots)):
        with kb.Solver(nx, ny, prec, p) as s:
            s.set_kernel("fused_dense")
            s.reset_disk(cx, cy, 6.0)
            done = 0
            for mk in marks:
                s.step(mk - done)
                done = mk
                for name, x, y in zip(("phi", "T", "theta"), got[mk][m], s.fields()):
                    assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), (m, mk, name)
    assert stats["listed_launches"] >= 350 and stats["stale_answers"] == 0, stats


def test_chunk_shape_changes_mid_run(kb):
    """The chunk shape changed between steps, in both launch forms, leaves
    every member bit for bit on a plain handle's trajectory; an unknown
    shape is refused and changes nothing."""
    sizes, params = members(kb)
    init = initial(sizes, "f32")
    with kb.Batch(sizes, params, "f32") as b:
        for m, st in enumerate(init):
            b.set_fields(m, *st)
        with pytest.raises(kb.KGError):
            b.set_chunk_shape(3, 3)
        assert b.chunk_shape == (2, 2)
        done = 0
        for shape, mode, g, n in [((8, 1), "launch", 0, 5), ((4, 2), "graph", 2, 4),
                                  ((2, 4), "graph", 2, 6), ((2, 2), "launch", 0, 3)]:
            b.set_chunk_shape(*shape)
            b.set_mode(mode, g)
            b.step(n)
            done += n
        for m, (nx, ny) in enumerate(sizes):
            ref = plain(kb, nx, ny, "f32", params[m], init[m], done)
            for x, y in zip(b.fields(m), ref):
                assert np.array_equal(x, y), m


@pytest.mark.parametrize("prec,shape,g", [("f32", (2, 2), 2), ("f32", (8, 1), 8),
                                          ("f64", (2, 2), 8)])
def test_graph_mode_device_lists_match_plain_solver_bitwise(kb, prec, shape, g):
    """KG_BATCH_GRAPH_LISTS over a batch of 256+ chunks keeps its chunk
    lists on the device (request and list kernels driven by the device step
    counter, no host wait): replays of a captured graph, and the steps
    between whole replays, leave every member bit for bit on a dense plain
    handle's trajectory, and the lists are used."""
    sizes, params, spots = list_batch(kb)
    marks = [1, 9, 40, 151, 400]
    with kb.Batch(sizes, params, prec) as b:
        b.set_chunk_shape(*shape)
        b.set_mode("graph_lists", g)
        assert b.mode() == ("graph_lists", g)
        for m, (cx, cy) in enumerate(spots):
            b.reset_disk(m, cx, cy, 6.0)
        got, done = {}, 0
        for mk in marks:
            b.step(mk - done)
            done = mk
            got[mk] = [b.fields(m) for m in range(len(sizes))]
        stats = b.skip_stats()
    check_against_dense(kb, prec, sizes, params, spots, got, marks)
    # one list per 7 steps from step 7 on, each counted as 7 launches
    assert stats["listed_launches"] >= 350 and stats["stale_answers"] == 0, stats


def test_lists_across_mode_and_shape_changes(kb):
    """Host lists (launch mode), device lists and no lists (the graph
    modes) in turn, with chunk-shape changes, mid-period switches and
    replays of different lengths: every member stays bit for bit on a dense
    plain handle's trajectory."""
    sizes, params, spots = list_batch(kb)
    plan = [("launch", 0, (2, 2), 23), ("graph_lists", 2, (2, 2), 30),
            ("graph_lists", 8, (4, 1), 41), ("graph", 4, (4, 1), 17), ("launch", 0, (4, 1), 12),
            ("graph_lists", 32, (2, 4), 70), ("graph", 8, (2, 4), 16),
            ("graph_lists", 8, (2, 4), 29), ("launch", 0, (2, 2), 9)]
    marks, done = [], 0
    with kb.Batch(sizes, params, "f32") as b:
        for m, (cx, cy) in enumerate(spots):
            b.reset_disk(m, cx, cy, 6.0)
        got = {}
        for mode, g, shape, n in plan:
            b.set_chunk_shape(*shape)
            b.set_mode(mode, g)
            b.step(n)
            done += n
            marks.append(done)
            got[done] = [b.fields(m) for m in range(len(sizes))]
    check_against_dense(kb, "f32", sizes, params, spots, got, marks)


@pytest.mark.parametrize("mode,g", [("launch", 0), ("graph", 4)])
def test_uniform_boundary_batches_and_a_boundary_change(kb, mode, g):
    """A batch whose members all share one boundary runs the step kernel
    compiled for that boundary alone; a member switched to the other
    boundary mid-run (a captured graph must then be rebuilt) keeps every
    member bit for bit on its plain handle's trajectory."""
    sizes = [(96, 64), (70, 90), (128, 40)]
    params = [kb.default_params(delta=0.03 + 0.01 * k, aniso=4.0 + k, noise_a=0.01, boundary=0,
                                seed=10 + k, instance=k) for k in range(3)]
    init = initial(sizes, "f32")
    with kb.Batch(sizes, params, "f32") as b:
        b.set_mode(mode, g)
        for m, st in enumerate(init):
            b.set_fields(m, *st)
        b.step(8)
        changed = params[1].replace(boundary=1)
        b.set_params(1, changed)
        b.step(8)
        got = [b.fields(m) for m in range(len(sizes))]
    for m, (nx, ny) in enumerate(sizes):
        with kb.Solver(nx, ny, "f32", params[m]) as s:
            s.set_fields(*init[m])
            s.step(8)
            if m == 1:
                s.params = changed
            s.step(8)
            for x, y in zip(got[m], s.fields()):
                assert np.array_equal(x, y), m


@pytest.mark.parametrize("mode,g", [("graph", 8), ("graph_lists", 2)])
def test_one_member_batch_is_a_replayable_single_grid(kb, mode, g):
    """A batch of one member is the graph-replayable form of a single grid
    (kg_step refuses stream capture): replays of its captured steps match a
    plain handle bit for bit, including a step count that is not a whole
    number of replays."""
    nx, ny = 640, 480
    p = kb.default_params(delta=0.04, aniso=6.0, K=2.0, theta0=0.2, noise_a=0.01, boundary=1,
                          seed=3)
    with kb.Batch([(nx, ny)], [p], "f32") as b:
        b.set_mode(mode, g)
        b.reset_seeds(4, 5.0, 8, 11, member=0)
        b.step(45)
        got = b.fields(0)
    with kb.Solver(nx, ny, "f32", p) as s:
        s.reset_seeds(4, 5.0, 8, 11)
        s.step(45)
        ref = s.fields()
    for x, y in zip(got, ref):
        assert np.array_equal(x, y)
