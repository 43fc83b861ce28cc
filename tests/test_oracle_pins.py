"""Pins of the CPU oracle (oracle/kobayashi_oracle.c) -- CPU only.

The reference (src/Kobayashi.cpp) cannot be built here: its header includes
the un-vendored DXViewer's Win32App.h.  It holds no tests, fixtures or golden
vectors either.  So every function of the oracle is pinned here by a check
that the mathematics fixes, chosen so that a plausible transcription mistake
(a dropped term, a wrong sign, a swapped gx/gy, /2dx instead of /dx, a wrong
stencil weight or boundary map) fails at least one of them:

  function (reference lines)          pinned by
  Philox4x32-10 noise source          published known-answer vectors
  9-point Laplacians (:142-151)       Fourier / DCT eigen-symbols, Allen-Cahn vs scipy
  reaction + m(T) (:206, :211-214)    uniform-field closed form
  T update with new phi (:215)        uniform-field closed form, sum(T-K phi) invariant
  theta branches (:154-167)           truth table against atan2
  eps, eps', t1, t2, t3 (:170-204)    analytic continuum operator, 2nd-order convergence;
                                      rotation / reflection equivariance
  periodic / zero-flux maps           translation invariance, locality, DCT symbols
  initial states (:98-123)            explicit cell sets
"""
import numpy as np
import pytest
from scipy import ndimage

pytestmark = []

FLT_EPS = float(np.finfo(np.float32).eps)


# ---------------------------------------------------------------- Philox
def test_philox_known_answers(O):
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "philox4x32_10_kat.txt")
    rows = [l.split() for l in open(path) if l.strip() and not l.startswith("#")]
    assert len(rows) == 3
    for r in rows:
        v = [int(x, 16) for x in r]
        out = O.philox4x32_10(v[0:4], v[4:6])
        assert out.tolist() == v[6:10]


def test_noise_range_and_moments(O):
    p = O.default_params(seed=12345, instance=3)
    chi = np.array([O.noise(p, i, j, 7) for j in range(64) for i in range(64)])
    assert chi.min() >= -0.5 and chi.max() < 0.5
    assert abs(chi.mean()) < 0.02
    assert abs(chi.var() - 1.0 / 12.0) < 0.01
    # 24-bit dyadic values: exactly representable in float32
    assert np.all(chi.astype(np.float32).astype(np.float64) == chi)
    # counter parts: step, instance and seed all change the sample
    assert O.noise(p, 5, 9, 7) != O.noise(p, 5, 9, 8)
    assert O.noise(p, 5, 9, 7) != O.noise(p.replace(instance=4), 5, 9, 7)
    assert O.noise(p, 5, 9, 7) != O.noise(p.replace(seed=1), 5, 9, 7)


# ------------------------------------------------------- fixed points
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("phi0", [0.0, 1.0])
@pytest.mark.parametrize("boundary", [0, 1])
def test_pure_phase_is_fixed_point(O, dtype, phi0, boundary):
    p = O.default_params(boundary=boundary, noise_a=0.01, seed=3)
    s = O.zeros(24, 20, dtype)
    s.phi[:] = phi0
    s.T[:] = 0.25
    s.theta[:] = np.linspace(0, 6, s.theta.size).reshape(s.theta.shape)
    out = O.run(s, p, 5)
    assert np.array_equal(out.phi, s.phi)
    assert np.array_equal(out.T, s.T)
    assert np.array_equal(out.theta, s.theta)   # flat gradient: theta never rewritten


# ------------------------------------------------ uniform closed form
@pytest.mark.parametrize("p0,t0", [(0.25, 0.0), (0.75, 0.5), (0.5, 1.25), (0.125, -0.5)])
def test_uniform_field_reaction_closed_form(O, p0, t0):
    """phi, T uniform (dyadic, so every Laplacian sum is exact and zero):
    phi' = p + p(1-p)(p - 1/2 + (alpha/pi) atan(gamma (Teq - t))) dt / tau,
    T'   = t + K (phi' - p)            (Kobayashi 1993, eqs. for p and T)."""
    p = O.default_params(delta=0.04, aniso=5.0, K=1.8)
    s = O.zeros(16, 12, np.float64)
    s.phi[:] = p0
    s.T[:] = t0
    out = O.step(s, p)
    m = p.alpha / np.pi * np.arctan(p.gamma * (p.teq - t0))
    phi1 = p0 + p0 * (1 - p0) * (p0 - 0.5 + m) * p.dt / p.tau
    t1 = t0 + p.K * (phi1 - p0)
    np.testing.assert_allclose(out.phi, phi1, rtol=1e-14, atol=0)
    np.testing.assert_allclose(out.T, t1, rtol=1e-13, atol=1e-15)


def test_noise_enters_double_well(O):
    """With noise the uniform-field update gains a p(1-p) chi(i,j,step) dt/tau."""
    p = O.default_params(noise_a=0.02, seed=99, instance=2)
    p0, t0 = 0.25, 0.5
    s = O.zeros(16, 8, np.float64)
    s.phi[:] = p0
    s.T[:] = t0
    out = O.step(s, p, step_index=11)
    m = p.alpha / np.pi * np.arctan(p.gamma * (p.teq - t0))
    chi = np.array([[O.noise(p, i, j, 11) for i in range(16)] for j in range(8)])
    phi1 = p0 + p0 * (1 - p0) * (p0 - 0.5 + m + p.noise_a * chi) * p.dt / p.tau
    np.testing.assert_allclose(out.phi, phi1, rtol=1e-14, atol=0)


# --------------------------------------------------- Laplacian symbols
def _symbol(a, b, dx):
    return (4 * np.cos(a) + 4 * np.cos(b) + 4 * np.cos(a) * np.cos(b) - 12) / (3 * dx * dx)


@pytest.mark.parametrize("k,l", [(1, 0), (2, 3), (5, 1), (7, 7)])
def test_heat_step_periodic_fourier_symbol(O, k, l):
    """phi = 0: T' = T + dt L9 T; cos(a i) cos(b j) is an eigenvector of the
    periodic 9-point Laplacian with symbol (4ca + 4cb + 4ca cb - 12)/(3dx^2)."""
    nx, ny = 32, 24
    p = O.default_params()
    a, b = 2 * np.pi * k / nx, 2 * np.pi * l / ny
    i = np.arange(nx)[None, :]
    j = np.arange(ny)[:, None]
    s = O.zeros(nx, ny, np.float64)
    s.T[:] = 0.3 * np.cos(a * i) * np.cos(b * j)
    out = O.step(s, p)
    expect = s.T * (1 + p.dt * _symbol(a, b, p.dx))
    np.testing.assert_allclose(out.T, expect, rtol=0, atol=1e-13)
    assert np.array_equal(out.phi, s.phi)


@pytest.mark.parametrize("k,l", [(1, 0), (2, 3), (6, 5)])
def test_heat_step_zero_flux_dct_symbol(O, k, l):
    """Zero-flux clamps indices (ghost = boundary cell); the DCT-II products
    cos(pi k (i+1/2)/n) cos(pi l (j+1/2)/m) are then eigenvectors."""
    nx, ny = 20, 28
    p = O.default_params(boundary=1)
    a, b = np.pi * k / nx, np.pi * l / ny
    i = np.arange(nx)[None, :]
    j = np.arange(ny)[:, None]
    s = O.zeros(nx, ny, np.float64)
    s.T[:] = 0.3 * np.cos(a * (i + 0.5)) * np.cos(b * (j + 0.5))
    out = O.step(s, p)
    expect = s.T * (1 + p.dt * _symbol(a, b, p.dx))
    np.testing.assert_allclose(out.T, expect, rtol=0, atol=1e-13)


# ------------------------------------------------ isotropic Allen-Cahn
@pytest.mark.parametrize("boundary,mode", [(0, "wrap"), (1, "nearest")])
def test_delta_zero_is_allen_cahn(O, boundary, mode):
    """delta = 0: eps == eps_bar, eps' == 0, every cross term vanishes and the
    step is forward-Euler Allen-Cahn with the 9-point Laplacian."""
    rng = np.random.default_rng(0)
    nx, ny = 30, 22
    p = O.default_params(delta=0.0, boundary=boundary)
    s = O.zeros(nx, ny, np.float64)
    s.phi[:] = ndimage.uniform_filter(rng.random((ny, nx)), 3, mode="wrap")
    s.T[:] = rng.random((ny, nx)) * 0.8
    s.theta[:] = rng.random((ny, nx)) * 6
    out = O.step(s, p)
    w = np.array([[1, 2, 1], [2, -12, 2], [1, 2, 1]], float) / (3 * p.dx * p.dx)
    lap = ndimage.convolve(s.phi, w, mode=mode)
    lapT = ndimage.convolve(s.T, w, mode=mode)
    m = p.alpha / np.pi * np.arctan(p.gamma * (p.teq - s.T))
    phi1 = s.phi + (p.eps_bar ** 2 * lap + s.phi * (1 - s.phi) * (s.phi - 0.5 + m)) * p.dt / p.tau
    np.testing.assert_allclose(out.phi, phi1, rtol=0, atol=1e-12)
    np.testing.assert_allclose(out.T, s.T + lapT * p.dt + p.K * (phi1 - s.phi), rtol=0, atol=1e-12)


# ------------------------------------------------------- theta branches
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_theta_truth_table(O, dtype):
    """theta is rewritten only when the gradient is 'non-flat'
    (src/Kobayashi.cpp:154-167) and then equals atan2(gy, gx) in the branch's
    range; otherwise the previous theta persists."""
    rng = np.random.default_rng(1)
    nx, ny = 48, 40
    levels = np.array([0.0, 0.0, 0.5, 1.0, 1e-9, 3e-9, 0.25], dtype)
    s = O.zeros(nx, ny, dtype)
    s.phi[:] = levels[rng.integers(0, len(levels), (ny, nx))]
    s.theta[:] = rng.uniform(-2, 8, (ny, nx))
    q = O.pass1(s, O.default_params())
    gx, gy, th = q["gx"].astype(float), q["gy"].astype(float), q["theta"].astype(float)
    old = s.theta.astype(float)
    E = FLT_EPS
    flatx = np.abs(gx) <= E
    classes = {
        "flat": flatx & (np.abs(gy) <= E),
        "down": flatx & (gy < -E),
        "up": flatx & (gy > E),
        "q4": (gx > E) & (gy < -E),
        "q1": (gx > E) & (gy > E),
        "xonly": (gx > E) & (np.abs(gy) <= E),
        "left": gx < -E,
    }
    for name, mask in classes.items():
        assert mask.sum() > 5, name          # every branch is exercised
    tol = 4e-7 if dtype == np.float32 else 1e-14
    assert np.array_equal(th[classes["flat"]], old[classes["flat"]])
    assert np.array_equal(th[classes["xonly"]], old[classes["xonly"]])
    assert np.allclose(th[classes["down"]], -np.pi / 2, atol=tol)
    assert np.allclose(th[classes["up"]], np.pi / 2, atol=tol)
    fired = classes["q4"] | classes["q1"] | classes["left"]
    ref = np.mod(np.arctan2(gy, gx), 2 * np.pi)
    np.testing.assert_allclose(th[fired], ref[fired], rtol=0, atol=4 * tol)
    assert np.all((th[classes["q1"]] > 0) & (th[classes["q1"]] < np.pi / 2 + tol))
    assert np.all((th[classes["q4"]] > 1.5 * np.pi - tol) & (th[classes["q4"]] < 2 * np.pi + tol))
    assert np.all((th[classes["left"]] > np.pi / 2 - tol) & (th[classes["left"]] < 1.5 * np.pi + tol))


# ------------------------------------ analytic continuum of the operator
def _continuum_case(alpha):
    """phi = 0.5 + 0.35 (x cos a + y sin a) + 0.08 sin(3x + 1) cos(2y) on [0,1]^2:
    |grad phi| >= 0.35 - 0.24 - 0.16 > 0 everywhere, so theta is smooth."""
    ca, sa = np.cos(alpha), np.sin(alpha)

    def fields(x, y):
        s3, c3 = np.sin(3 * x + 1), np.cos(3 * x + 1)
        s2, c2 = np.sin(2 * y), np.cos(2 * y)
        phi = 0.5 + 0.35 * (x * ca + y * sa) + 0.08 * s3 * c2
        px = 0.35 * ca + 0.24 * c3 * c2
        py = 0.35 * sa - 0.16 * s3 * s2
        pxx = -0.72 * s3 * c2
        pyy = -0.32 * s3 * c2
        pxy = -0.48 * c3 * s2
        T = 0.2 + 0.1 * x * x - 0.05 * y + 0.3 * np.sin(2 * x + y)
        lapT = 0.2 - 1.5 * np.sin(2 * x + y)
        return phi, px, py, pxx, pyy, pxy, T, lapT
    return fields


@pytest.mark.parametrize("alpha", [0.3, 2.0, 3.6, 5.3])
@pytest.mark.parametrize("aniso,theta0", [(6.0, 0.0), (4.0, 0.2), (5.0, 1.1)])
def test_operator_converges_to_continuum(O, alpha, aniso, theta0):
    """(phi' - phi) tau / dt of one step converges at second order to
        4 [ -d_x(e e' phi_y) + d_y(e e' phi_x) + grad(e^2).grad(phi) ] + (4/3) e^2 lap(phi) + reaction
    where e = eps(theta), theta = atan2(phi_y, phi_x).  Both factors are the
    reference's: its first differences divide by dx, not 2dx (src/Kobayashi.cpp:139-140,
    :190-203), so each of the two nested first derivatives carries a factor 2;
    its 9-point stencil (2 axis + diagonals - 12 centre)/(3 dx^2) (:142-151) has
    second moment 8 dx^2, i.e. it converges to (4/3) lap.
    Expanding analytically (e'' = -eps_bar delta j^2 cos):
        -d_x(e e' phi_y) + d_y(e e' phi_x) = (e'^2 + e e'')(th_y phi_x - th_x phi_y)
        grad(e^2).grad(phi)                = 2 e e' (th_x phi_x + th_y phi_y)
    with th_x = (phi_x phi_xy - phi_y phi_xx)/|g|^2, th_y = (phi_x phi_yy - phi_y phi_xy)/|g|^2.
    T: (T' - T - K (phi' - phi)) / dt -> (4/3) lap T."""
    fields = _continuum_case(alpha)
    ebar, delta = 0.1, 0.2
    errs, errsT, scale = [], [], None
    for n in (48, 96, 192):
        dx = 1.0 / n
        p = O.default_params(dx=dx, dy=dx, dt=1e-3, tau=1.0, eps_bar=ebar, delta=delta,
                             aniso=aniso, theta0=theta0, boundary=1)
        xc = (np.arange(n) + 0.5) * dx
        X, Y = np.meshgrid(xc, xc)
        phi, px, py, pxx, pyy, pxy, T, lapT = fields(X, Y)
        s = O.State(phi.copy(), T.copy(), np.zeros_like(phi))
        out = O.step(s, p)
        disc = (out.phi - phi) * p.tau / p.dt
        g2 = px * px + py * py
        th = np.arctan2(py, px)
        thx = (px * pxy - py * pxx) / g2
        thy = (px * pyy - py * pxy) / g2
        arg = aniso * (th - theta0)
        e = ebar * (1 + delta * np.cos(arg))
        e1 = -ebar * aniso * delta * np.sin(arg)
        e2 = -ebar * aniso * aniso * delta * np.cos(arg)
        A = (e1 * e1 + e * e2) * (thy * px - thx * py) + 2 * e * e1 * (thx * px + thy * py)
        m = p.alpha / np.pi * np.arctan(p.gamma * (p.teq - T))
        cont = 4 * A + 4.0 / 3.0 * e * e * (pxx + pyy) + phi * (1 - phi) * (phi - 0.5 + m)
        sl = (slice(3, n - 3), slice(3, n - 3))
        errs.append(np.abs(disc - cont)[sl].max())
        discT = (out.T - T - p.K * (out.phi - phi)) / p.dt
        errsT.append(np.abs(discT - 4.0 / 3.0 * lapT)[sl].max())
        scale = np.abs(4 * A)[sl].max()
    assert scale > 1e-3
    assert errs[1] / errs[0] < 0.32 and errs[2] / errs[1] < 0.3, errs
    assert errs[2] < 2e-3 * scale, (errs, scale)
    assert errsT[1] / errsT[0] < 0.32 and errsT[2] / errsT[1] < 0.3, errsT


# --------------------------------------------------------- symmetries
def _smooth_periodic(rng, n, amp=0.3):
    i = np.arange(n)
    X, Y = np.meshgrid(i, i)
    f = np.full((n, n), 0.5)
    for _ in range(6):
        k, l = rng.integers(-3, 4, 2)
        ph = rng.uniform(0, 2 * np.pi)
        f += amp / 6 * np.cos(2 * np.pi * (k * X + l * Y) / n + ph)
    return f


def test_rotation_equivariance_j4(O):
    """j = 4 makes eps(theta + pi/2) == eps(theta): the step commutes with a
    90-degree rotation of a square periodic grid (swapping the roles of
    t1/t2 and gx/gy)."""
    rng = np.random.default_rng(4)
    n = 32
    p = O.default_params(aniso=4.0, delta=0.05, eps_bar=0.02)
    s = O.State(_smooth_periodic(rng, n), _smooth_periodic(rng, n), np.zeros((n, n)))
    ref = O.step(s, p)
    for k in (1, 3):
        r = O.State(*(np.ascontiguousarray(np.rot90(a, k)) for a in (s.phi, s.T, s.theta)))
        out = O.step(r, p)
        np.testing.assert_allclose(out.phi, np.rot90(ref.phi, k), rtol=0, atol=1e-13)
        np.testing.assert_allclose(out.T, np.rot90(ref.T, k), rtol=0, atol=1e-13)


@pytest.mark.parametrize("axis", [0, 1])
def test_reflection_equivariance_even_j(O, axis):
    rng = np.random.default_rng(5)
    n = 28
    p = O.default_params(aniso=6.0, delta=0.05, eps_bar=0.02)
    s = O.State(_smooth_periodic(rng, n), _smooth_periodic(rng, n), np.zeros((n, n)))
    ref = O.step(s, p)
    r = O.State(*(np.ascontiguousarray(np.flip(a, axis)) for a in (s.phi, s.T, s.theta)))
    out = O.step(r, p)
    np.testing.assert_allclose(out.phi, np.flip(ref.phi, axis), rtol=0, atol=1e-13)
    np.testing.assert_allclose(out.T, np.flip(ref.T, axis), rtol=0, atol=1e-13)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_periodic_translation_invariance_bitwise(O, dtype):
    rng = np.random.default_rng(6)
    nx, ny = 36, 26
    p = O.default_params()
    s = O.State(rng.random((ny, nx)), rng.random((ny, nx)), rng.uniform(0, 6, (ny, nx))).astype(dtype)
    ref = O.run(s, p, 3)
    sh = (5, 11)
    r = O.State(*(np.roll(a, sh, (0, 1)) for a in (s.phi, s.T, s.theta)))
    out = O.run(r, p, 3)
    for a, b in ((out.phi, ref.phi), (out.T, ref.T), (out.theta, ref.theta)):
        assert np.array_equal(a, np.roll(b, sh, (0, 1)))


def test_zero_flux_only_touches_boundary_band(O):
    """One step: periodic and zero-flux agree bitwise on cells >= 2 from
    every edge (stencil radius 2: pass 2 reads pass-1 values of neighbours)."""
    rng = np.random.default_rng(7)
    nx, ny = 30, 24
    s = O.State(rng.random((ny, nx)), rng.random((ny, nx)), rng.uniform(0, 6, (ny, nx)))
    a = O.step(s, O.default_params(boundary=0))
    b = O.step(s, O.default_params(boundary=1))
    inner = (slice(2, ny - 2), slice(2, nx - 2))
    assert np.array_equal(a.phi[inner], b.phi[inner])
    assert np.array_equal(a.T[inner], b.T[inner])
    assert not np.array_equal(a.phi, b.phi)


@pytest.mark.parametrize("boundary,pad", [(0, "wrap"), (1, "edge")])
def test_pass2_boundary_map_as_ghost_cells(O, boundary, pad):
    """Pass 2 reads eps^2, eps eps' gx and eps eps' gy of the four neighbours.
    Rebuild it from the oracle's own pass-1 arrays with ghost cells made by
    array padding -- 'wrap' for periodic (src/Kobayashi.cpp:184-187), 'edge'
    (replicate the boundary cell's pass-1 values) for zero-flux -- and compare
    with the oracle's step, including every boundary cell.  This pins the
    index map of pass 2 separately from pass 1's (pinned by the DCT symbols
    and the Allen-Cahn check)."""
    rng = np.random.default_rng(10)
    nx, ny = 22, 18
    p = O.default_params(boundary=boundary, aniso=5.0, delta=0.05, eps_bar=0.02, theta0=0.3)
    s = O.State(rng.random((ny, nx)), rng.random((ny, nx)) * 0.7, rng.uniform(0, 6, (ny, nx)))
    q = O.pass1(s, p)
    e, ed, gx, gy = q["eps"], q["eps_deriv"], q["gx"], q["gy"]
    P = lambda a: np.pad(a, 1, mode=pad)
    e2, A, B = P(e * e), P(e * ed * gx), P(e * ed * gy)
    N = lambda a: a[2:, 1:-1]
    S = lambda a: a[:-2, 1:-1]
    E = lambda a: a[1:-1, 2:]
    W = lambda a: a[1:-1, :-2]
    t1 = (N(A) - S(A)) / p.dy
    t2 = -(E(B) - W(B)) / p.dx
    t3 = (E(e2) - W(e2)) / p.dx * gx + (N(e2) - S(e2)) / p.dy * gy
    m = p.alpha / np.pi * np.arctan(p.gamma * (p.teq - s.T))
    phi1 = s.phi + (t1 + t2 + e * e * q["lap_phi"] + t3
                    + s.phi * (1 - s.phi) * (s.phi - 0.5 + m)) * p.dt / p.tau
    out = O.step(s, p)
    np.testing.assert_allclose(out.phi, phi1, rtol=0, atol=1e-13)
    np.testing.assert_allclose(out.T, s.T + q["lap_T"] * p.dt + p.K * (phi1 - s.phi), rtol=0, atol=1e-13)


@pytest.mark.parametrize("axis", [0, 1])
def test_zero_flux_mirror_symmetry(O, axis):
    """A state symmetric about the domain centre stays symmetric under the
    zero-flux map (j even, theta0 = 0)."""
    rng = np.random.default_rng(8)
    n = 24
    f = rng.random((n, n))
    g = rng.random((n, n))
    f = 0.5 * (f + np.flip(f, axis))
    g = 0.5 * (g + np.flip(g, axis))
    s = O.State(f, g, np.zeros((n, n)))
    out = O.step(s, O.default_params(boundary=1, aniso=6.0))
    np.testing.assert_allclose(out.phi, np.flip(out.phi, axis), rtol=0, atol=1e-13)
    np.testing.assert_allclose(out.T, np.flip(out.T, axis), rtol=0, atol=1e-13)


# -------------------------------------------------------- invariants
@pytest.mark.parametrize("boundary", [0, 1])
@pytest.mark.parametrize("noise_a", [0.0, 0.01])
def test_sum_T_minus_K_phi_invariant(O, boundary, noise_a):
    """T' - K phi' = T - K phi + dt L9 T and sum(L9 T) == 0 for both boundary
    maps, so sum(T - K phi) is invariant up to rounding."""
    n = 48
    p = O.default_params(boundary=boundary, noise_a=noise_a, seed=5, delta=0.04)
    s = O.init_disk(n, n, 20, 27, 6.0)
    out = O.run(s, p, 200)
    inv0 = (s.T - p.K * s.phi).sum()
    inv1 = (out.T - p.K * out.phi).sum()
    assert abs(inv1 - inv0) < 1e-11 * (np.abs(out.T).sum() + p.K * np.abs(out.phi).sum())
    assert out.phi.max() > 0.9 and abs(out.T).sum() > 1  # something did evolve


# --------------------------------------------------- precision budget
@pytest.mark.parametrize("case", ["c0", "c1"])
def test_fp32_fp64_budget_disk(O, case):
    """The reference's float arithmetic vs the FP64 oracle on disk seeds
    (no exact-symmetry theta flips): relative L2 error at 100 steps."""
    if case == "c0":
        p = O.default_params(delta=0.02, aniso=4.0, theta0=0.2, noise_a=0.01, boundary=1)
    else:
        p = O.default_params(delta=0.04, aniso=6.0, K=2.0, theta0=0.2, noise_a=0.01, boundary=1)
    s = O.init_disk(96, 96, 48, 48, 5.0)
    a = O.run(s.astype(np.float32), p, 100)
    b = O.run(s, p, 100)
    for x, y in ((a.phi, b.phi), (a.T, b.T)):
        assert np.linalg.norm(x - y) / np.linalg.norm(y) < 1e-6


# ---------------------------------------------------- initial states
def test_init_plus_matches_reference_nucleus(O):
    """_vectorInit + _createNucleus(nx/2, ny/2): a 5-cell plus of phi = 1."""
    for nx, ny in ((250, 250), (31, 17)):
        s = O.init_plus(nx, ny)
        cx, cy = nx // 2, ny // 2
        cells = {(cx, cy), (cx - 1, cy), (cx + 1, cy), (cx, cy - 1), (cx, cy + 1)}
        ys, xs = np.nonzero(s.phi)
        assert set(zip(xs.tolist(), ys.tolist())) == cells
        assert np.all(s.phi[ys, xs] == 1.0)
        assert not s.T.any() and not s.theta.any()


def test_init_disk_and_seeds(O):
    s = O.init_disk(40, 30, 12, 17, 5.0)
    ys, xs = np.nonzero(s.phi)
    assert np.all((xs - 12) ** 2 + (ys - 17) ** 2 <= 25)
    assert len(xs) == 81           # lattice points with i^2 + j^2 <= 25
    s, cxy = O.init_seeds(200, 150, 9, 5.0, 8, 1234)
    assert cxy.shape == (9, 2)
    assert np.all(cxy[:, 0] >= 8) and np.all(cxy[:, 0] < 192)
    assert np.all(cxy[:, 1] >= 8) and np.all(cxy[:, 1] < 142)
    s2, cxy2 = O.init_seeds(200, 150, 9, 5.0, 8, 1234)
    assert np.array_equal(cxy, cxy2) and np.array_equal(s.phi, s2.phi)
    _, cxy3 = O.init_seeds(200, 150, 9, 5.0, 8, 1235)
    assert not np.array_equal(cxy, cxy3)
