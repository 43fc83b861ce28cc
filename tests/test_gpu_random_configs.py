"""Randomised configurations against the CPU oracle (needs a GPU).

Sixty-four configurations drawn from a fixed seed: grid sizes from 3 x 3
to 259 x 199 (widths that are and are not multiples of 4, ragged last tiles),
every parameter of the model drawn over and beyond the reference GUI's
slider ranges (negative K, non-integer anisotropy, theta0 outside
[0, 2 pi), noise on and off), both boundaries, a disk or random fields, a
random starting step count, 1 to 30 steps, through the default kernel (tile
status, skips) and the dense one.  The same bars as tests/test_gpu_parity.py:

  FP64 kernel vs FP64 oracle: max |diff| <= 1e-10 on phi and T; theta equal
      to 1e-10 modulo 2 pi where the stencil differences are >= 1e-3.
  FP32 kernel vs FP32 oracle after one step: max |diff| <= 2e-6.

Twenty-four more drawn configurations form one FP64 batch (members of
every drawn size, parameter set and boundary in one launch per step), each
member held to the same 1e-10 against its own oracle run.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N_CONFIGS = 64


@pytest.fixture(scope="module")
def kb():
    import crystalgrowth_amd as kb
    if kb.device_count() < 1:
        pytest.fail("no GPU visible")
    return kb


def draw(kb, O, k):
    """Configuration k: (nx, ny, oracle params, initial state, steps, step0)."""
    rng = np.random.default_rng(20261203 + 7919 * k)
    while True:
        nx, ny = int(rng.integers(3, 260)), int(rng.integers(3, 200))
        kw = dict(tau=float(rng.uniform(2.5e-4, 6e-4)), eps_bar=float(rng.uniform(0.005, 0.014)),
                  K=float(rng.uniform(-0.5, 2.5)), delta=float(rng.uniform(0.0, 0.08)),
                  aniso=float(rng.choice([2.0, 4.0, 4.5, 5.0, 6.0, 8.0])),
                  alpha=float(rng.uniform(0.1, 1.0)), gamma=float(rng.uniform(1.0, 20.0)),
                  teq=float(rng.uniform(0.5, 1.5)), theta0=float(rng.uniform(-1.0, 7.0)),
                  noise_a=float(rng.choice([0.0, 0.01, 0.05])),
                  boundary=int(rng.integers(0, 2)), seed=int(rng.integers(0, 2 ** 63)),
                  instance=int(rng.integers(0, 1000)))
        p = O.default_params(**kw)
        try:
            kb.validate(nx, ny, kb.Params(**p.as_dict()))
        except Exception:
            continue
        break
    if rng.random() < 0.5 and min(nx, ny) >= 12:
        state = O.init_disk(nx, ny, int(rng.integers(2, nx - 2)), int(rng.integers(2, ny - 2)),
                            float(rng.uniform(2.0, 0.4 * min(nx, ny))))
    else:
        state = O.State(rng.random((ny, nx)), rng.random((ny, nx)) * 0.6,
                        rng.uniform(-1.0, 7.0, (ny, nx)))
    return nx, ny, p, state, int(rng.integers(1, 31)), int(rng.integers(0, 1 << 30))


def run_gpu(kb, nx, ny, p, state, nsteps, step0, precision, kernel):
    dt = np.float64 if precision == "f64" else np.float32
    with kb.Solver(nx, ny, precision, kb.Params(**p.as_dict())) as s:
        s.set_kernel(kernel)
        s.set_fields(state.phi.astype(dt), state.T.astype(dt), state.theta.astype(dt))
        s.step_count = step0
        s.step(nsteps)
        return s.fields()


def well_conditioned(phi, boundary):
    mode = "wrap" if boundary == 0 else "edge"
    q = np.pad(phi, 1, mode=mode)
    gx = np.abs(q[1:-1, 2:] - q[1:-1, :-2])
    gy = np.abs(q[2:, 1:-1] - q[:-2, 1:-1])
    return np.maximum(gx, gy) >= 1e-3


@pytest.mark.parametrize("kernel", ["fused", "fused_dense"])
@pytest.mark.parametrize("k", range(N_CONFIGS))
def test_fp64_random_config_matches_oracle(kb, O, k, kernel):
    nx, ny, p, state, nsteps, step0 = draw(kb, O, k)
    print(f"config {k}: {nx}x{ny} boundary {p.boundary} K {p.K:.3f} aniso {p.aniso} "
          f"noise {p.noise_a} steps {nsteps} from {step0}", flush=True)
    ref = O.run(state.astype(np.float64), p, nsteps, step0)
    phi, T, th = run_gpu(kb, nx, ny, p, state, nsteps, step0, "f64", kernel)
    assert np.max(np.abs(phi - ref.phi)) <= 1e-10
    assert np.max(np.abs(T - ref.T)) <= 1e-10
    # the angle is compared where the phase differences make it well defined
    # (the last step's input phase decides it)
    prev = O.run(state.astype(np.float64), p, nsteps - 1, step0) if nsteps > 1 else state
    m = well_conditioned(prev.phi.astype(np.float64), p.boundary)
    d = np.mod(th[m] - ref.theta[m] + np.pi, 2 * np.pi) - np.pi
    assert d.size == 0 or np.max(np.abs(d)) <= 1e-10


@pytest.mark.parametrize("k", range(N_CONFIGS))
def test_fp32_random_config_one_step_matches_oracle(kb, O, k):
    nx, ny, p, state, _, step0 = draw(kb, O, k)
    p32 = O.Params(**{name: (float(np.float32(v)) if isinstance(v, float) else v)
                      for name, v in p.as_dict().items()})
    s32 = state.astype(np.float32)
    ref = O.run(s32, p32, 1, step0)
    phi, T, _ = run_gpu(kb, nx, ny, p32, s32, 1, step0, "f32", "fused")
    assert np.max(np.abs(phi - ref.phi)) <= 2e-6
    assert np.max(np.abs(T - ref.T)) <= 2e-6


def test_fp64_random_batch_members_match_oracle(kb, O):
    members = [draw(kb, O, 500 + k) for k in range(24)]
    nsteps = 17
    with kb.Batch([(m[0], m[1]) for m in members], [kb.Params(**m[2].as_dict()) for m in members],
                  "f64") as b:
        for i, m in enumerate(members):
            st = m[3]
            b.set_fields(i, st.phi.astype(np.float64), st.T.astype(np.float64),
                         st.theta.astype(np.float64))
        b.step(nsteps)
        got = [b.fields(i) for i in range(len(members))]
    for i, m in enumerate(members):
        ref = O.run(m[3].astype(np.float64), m[2], nsteps, 0)
        assert np.max(np.abs(got[i][0] - ref.phi)) <= 1e-10, i
        assert np.max(np.abs(got[i][1] - ref.T)) <= 1e-10, i
