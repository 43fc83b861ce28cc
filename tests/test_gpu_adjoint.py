"""Adjoint (kg_adjoint) checked against finite differences (needs a GPU).

Tolerances: dL/d(parameter) of the FP64 adjoint vs central differences of the
FP64 forward with h = 1e-4 (relative to the parameter; absolute for theta0 and
the noise amplitude): <= 1e-3 relative.  The same vs central differences of
the FP64 CPU oracle on a tiny grid (the CPU reference gradient).  The FP32
adjoint vs the FP64 adjoint: <= 2e-2 relative.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NAMES = ("tau", "eps_bar", "K", "delta", "theta0", "alpha", "gamma", "teq", "noise_a")


@pytest.fixture(scope="module")
def kb():
    import crystalgrowth_amd as kb
    if kb.device_count() < 1:
        pytest.fail("no GPU visible")
    return kb


def setup_case(kb, nx, ny, boundary, nsteps):
    base = kb.default_params(delta=0.04, aniso=6.0, K=1.6, theta0=0.15, noise_a=0.01,
                             boundary=boundary, seed=5)
    yy, xx = np.mgrid[0:ny, 0:nx]
    phi0 = (((xx - nx // 2) ** 2 + (yy - ny // 2 + 1) ** 2) <= 30).astype(np.float64)
    T0 = 0.05 * np.cos(2 * np.pi * xx / nx) * np.sin(2 * np.pi * yy / ny)
    with kb.Solver(nx, ny, "f64", base.replace(delta=0.05, K=1.8)) as s:
        s.set_fields(phi0, T0)
        s.step(nsteps)
        target = s.phase()
    return base, phi0, T0, target


def loss_fwd(kb, base, phi0, T0, target, nsteps, prec="f64"):
    ny, nx = target.shape
    with kb.Solver(nx, ny, prec, base) as s:
        s.set_fields(phi0, T0)
        s.step(nsteps)
        d = s.phase().astype(np.float64) - target
        return float((d * d).sum())


def fd(kb, base, name, phi0, T0, target, nsteps):
    v = getattr(base, name)
    h = 1e-4 if name in ("theta0", "noise_a") else 1e-4 * abs(v)
    lp = loss_fwd(kb, base.replace(**{name: v + h}), phi0, T0, target, nsteps)
    lm = loss_fwd(kb, base.replace(**{name: v - h}), phi0, T0, target, nsteps)
    return (lp - lm) / (2 * h)


@pytest.mark.parametrize("boundary", [0, 1])
def test_parameter_gradients_match_central_differences(kb, boundary):
    nx, ny, n = 48, 40, 60
    base, phi0, T0, target = setup_case(kb, nx, ny, boundary, n)
    with kb.Solver(nx, ny, "f64", base) as s:
        s.set_fields(phi0, T0)
        g, _, _ = s.adjoint(n, target, checkpoint_every=7)
    assert g["loss"] > 0
    np.testing.assert_allclose(g["loss"], loss_fwd(kb, base, phi0, T0, target, n), rtol=1e-12)
    for name in NAMES:
        ref = fd(kb, base, name, phi0, T0, target, n)
        assert abs(g["d_" + name] - ref) <= 1e-3 * abs(ref) + 1e-12, (name, g["d_" + name], ref)


def test_field_gradients_match_central_differences(kb):
    """From a state without exact symmetry: on a symmetry axis of the disk gx
    is exactly 0, where the reference's angle is the constant pi/2 for
    |gx| <= FLT_EPSILON (derivative 0) while a finite perturbation of phi
    lands on the atan branch -- central differences cannot see the plateau.
    The step h = 1e-7 keeps the truncation error far below the tolerance at
    liquid cells (phi ~ 1e-3), whose loss is strongly curved: at cell (0, 0)
    the FP64 oracle's central differences read 7.3256e-5 (h = 1e-5),
    7.31798e-5 (1e-6) and 7.3179015e-5 (1e-7), converging to the adjoint."""
    nx, ny, n = 40, 36, 40
    base, phi0, T0, target = setup_case(kb, nx, ny, 1, n)
    rng = np.random.default_rng(3)
    phi0 = phi0 + 1e-3 * rng.random(phi0.shape)
    with kb.Solver(nx, ny, "f64", base) as s:
        s.set_fields(phi0, T0)
        _, dT, dP = s.adjoint(n, target, checkpoint_every=16, want_T0=True, want_phi0=True)
    rng = np.random.default_rng(0)
    cells = [(ny // 2, nx // 2 + 5), (0, 0), (ny - 1, 3), (ny // 2 + 6, nx // 2)]
    cells += [tuple(c) for c in rng.integers(0, [ny, nx], (4, 2))]
    h = 1e-7
    for (j, i) in cells:
        for field, grad in (("T", dT), ("phi", dP)):
            a, b = T0.copy(), phi0.copy()
            (a if field == "T" else b)[j, i] += h
            lp = loss_fwd(kb, base, b, a, target, n)
            a, b = T0.copy(), phi0.copy()
            (a if field == "T" else b)[j, i] -= h
            lm = loss_fwd(kb, base, b, a, target, n)
            ref = (lp - lm) / (2 * h)
            assert abs(grad[j, i] - ref) <= 1e-3 * abs(ref) + 1e-9, (field, j, i, grad[j, i], ref)


def test_cpu_reference_gradient(kb, O):
    """Central differences of the FP64 CPU oracle (the reference arithmetic)."""
    nx, ny, n = 24, 20, 30
    base, phi0, T0, target = setup_case(kb, nx, ny, 1, n)
    with kb.Solver(nx, ny, "f64", base) as s:
        s.set_fields(phi0, T0)
        g, _, _ = s.adjoint(n, target, checkpoint_every=n)
    op = O.Params(**base.as_dict())

    def oloss(p):
        st = O.run(O.State(phi0.copy(), T0.copy(), np.zeros_like(phi0)), p, n)
        d = st.phi - target
        return float((d * d).sum())

    for name in ("delta", "K", "eps_bar", "tau"):
        v = getattr(op, name)
        h = 1e-4 * abs(v)
        ref = (oloss(op.replace(**{name: v + h})) - oloss(op.replace(**{name: v - h}))) / (2 * h)
        assert abs(g["d_" + name] - ref) <= 1e-3 * abs(ref), (name, g["d_" + name], ref)


def test_checkpoint_interval_is_exact_and_state_is_final(kb):
    nx, ny, n = 40, 32, 45
    base, phi0, T0, target = setup_case(kb, nx, ny, 0, n)
    res = []
    for every in (1, 7, 45, 100):
        with kb.Solver(nx, ny, "f64", base) as s:
            s.set_fields(phi0, T0)
            g, dT, _ = s.adjoint(n, target, checkpoint_every=every, want_T0=True)
            assert s.step_count == n
            final = s.fields()
        res.append((g, dT))
    with kb.Solver(nx, ny, "f64", base) as s:
        s.set_fields(phi0, T0)
        s.step(n)
        for x, y in zip(final, s.fields()):
            assert np.array_equal(x, y)
    for g, dT in res[1:]:
        assert g == res[0][0]
        assert np.array_equal(dT, res[0][1])


def test_fp32_adjoint_close_to_fp64(kb):
    nx, ny, n = 64, 64, 100
    base, phi0, T0, target = setup_case(kb, nx, ny, 1, n)
    out = {}
    for prec in ("f32", "f64"):
        with kb.Solver(nx, ny, prec, base) as s:
            s.set_fields(phi0, T0)
            out[prec], _, _ = s.adjoint(n, target, checkpoint_every=25)
    for name in ("delta", "K", "eps_bar", "tau"):
        a, b = out["f32"]["d_" + name], out["f64"]["d_" + name]
        assert abs(a - b) <= 2e-2 * abs(b), (name, a, b)


def test_torch_autograd_function(kb):
    import torch
    from crystalgrowth_amd.autograd import kobayashi_loss
    nx, ny, n = 32, 32, 30
    base, phi0, T0, target = setup_case(kb, nx, ny, 1, n)
    p = torch.tensor([base.delta, base.K], dtype=torch.float64, requires_grad=True)
    t0 = torch.tensor(T0, requires_grad=True)
    L = kobayashi_loss(p, t0, names=("delta", "K"), base=base, phi0=phi0, theta=None,
                       target=target, nsteps=n, every=10, precision="f64")
    L.backward()
    with kb.Solver(nx, ny, "f64", base) as s:
        s.set_fields(phi0, T0)
        g, dT, _ = s.adjoint(n, target, 10, want_T0=True)
    assert L.item() == g["loss"]
    assert p.grad.tolist() == [g["d_delta"], g["d_K"]]
    assert np.array_equal(t0.grad.numpy(), dT)


def test_batched_adjoint_matches_single_handles(kb):
    """Every member of kg_batch_adjoint equals kg_adjoint of a plain handle
    with the same size, parameters and state, bit for bit."""
    sizes = [(48, 40), (64, 64), (37, 23)]
    cases = [setup_case(kb, nx, ny, b, 30) for (nx, ny), b in zip(sizes, (1, 0, 1))]
    params = [c[0].replace(instance=m) for m, c in enumerate(cases)]
    with kb.Batch(sizes, params, "f64") as b:
        for m, c in enumerate(cases):
            b.set_fields(m, c[1], c[2])
        gb, dTb = b.adjoint(30, [c[3] for c in cases], checkpoint_every=8, want_T0=True)
        assert b.step_count == 30
    for m, ((nx, ny), c) in enumerate(zip(sizes, cases)):
        with kb.Solver(nx, ny, "f64", params[m]) as s:
            s.set_fields(c[1], c[2])
            g, dT, _ = s.adjoint(30, c[3], checkpoint_every=8, want_T0=True)
        assert g == gb[m], m
        assert np.array_equal(dT, dTb[m])


def test_c3_fp32_adjoint_at_its_size_matches_fp64_central_differences(kb):
    """The adjoint configuration itself: 512^2 FP32, 1000 steps, checkpoint
    every 50, target from delta 0.04, K 1.8 (crystalgrowth_amd/configs.py c3).
    dL/d(delta, K, eps_bar, tau) and dL/dT0 at four cells (the three with
    the largest FP64 sensitivity and one off the front) of the FP32 adjoint
    against central differences (h = 1e-4, relative for the parameters,
    absolute for T0) of the FP64 forward: <= 1e-3 relative.  Measured on
    MI355X: <= 4e-5 for the parameters, <= 1.7e-5 for the T0 cells."""
    from crystalgrowth_amd import configs
    c = configs.C3
    n, steps, every = c["n"], c["steps"], c["every"]
    with kb.Solver(n, n, "f32", configs.c3_params("target")) as s:
        s.reset_disk(n // 2, n // 2, 5.0)
        s.step(steps)
        target = s.phase().astype(np.float64)
    base = configs.c3_params("guess")
    with kb.Solver(n, n, "f32", base) as s:
        s.reset_disk(n // 2, n // 2, 5.0)
        g32, dT32, _ = s.adjoint(steps, target.astype(np.float32), every, want_T0=True)
    with kb.Solver(n, n, "f64", base) as s:
        s.reset_disk(n // 2, n // 2, 5.0)
        phi0, T0, _ = s.fields()
        _, dT64, _ = s.adjoint(steps, target, every, want_T0=True)

    def loss(p, T0_):
        with kb.Solver(n, n, "f64", p) as s:
            s.set_fields(phi0, T0_)
            s.step(steps)
            d = s.phase() - target
            return float((d * d).sum())

    for name in ("delta", "K", "eps_bar", "tau"):
        v = getattr(base, name)
        h = 1e-4 * abs(v)
        ref = (loss(base.replace(**{name: v + h}), T0) - loss(base.replace(**{name: v - h}), T0)) / (2 * h)
        assert abs(g32["d_" + name] - ref) <= 1e-3 * abs(ref), (name, g32["d_" + name], ref)
    cells = list(np.argsort(np.abs(dT64).ravel())[::-1][:3]) + [n * (n // 2 + 30) + n // 2]
    h = 1e-4
    for k in cells:
        tp, tm = T0.copy(), T0.copy()
        tp.ravel()[k] += h
        tm.ravel()[k] -= h
        ref = (loss(base, tp) - loss(base, tm)) / (2 * h)
        got = float(dT32.ravel()[k])
        assert abs(got - ref) <= 1e-3 * abs(ref), (int(k), got, ref)
