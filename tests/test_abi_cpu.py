"""The C-ABI library without a GPU: it loads, exports every declared symbol,
and its host-only functions (parameters, validation, Philox) work."""
import ctypes
import os

import numpy as np
import pytest

from crystalgrowth_amd import _lib
from crystalgrowth_amd import kobayashi as kb


@pytest.fixture(scope="module")
def L():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail(f"{_lib.LIB_PATH} not built (run __graft_entry__.build())")
    return _lib.lib()


def test_exports_every_header_symbol(L):
    names = _lib.header_symbols()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.SIGNATURES), "binding and header disagree"


def test_abi_version(L):
    assert L.kg_abi_version() == 1


def test_default_params_are_reference_values(L):
    p = kb.default_params()
    f = lambda x: float(np.float32(x))
    assert (p.dx, p.dy, p.dt) == (f(0.03), f(0.03), f(1e-4))
    assert (p.tau, p.eps_bar, p.mu, p.K) == (f(3e-4), f(0.01), 1.0, f(1.6))
    assert (p.delta, p.aniso, p.alpha, p.gamma, p.teq) == (f(0.05), 6.0, f(0.9), 10.0, 1.0)
    assert (p.theta0, p.noise_a, p.boundary, p.seed) == (0.0, 0.0, 0, 0)


def test_default_params_match_oracle(L, O):
    a = kb.default_params().as_dict()
    b = O.default_params().as_dict()
    assert a == b


@pytest.mark.parametrize("bad,code", [
    (dict(dx=0.0), _lib.KG_ERR_INVALID_ARGUMENT),
    (dict(dt=float("nan")), _lib.KG_ERR_INVALID_ARGUMENT),
    (dict(tau=-1.0), _lib.KG_ERR_INVALID_ARGUMENT),
    (dict(boundary=7), _lib.KG_ERR_INVALID_ARGUMENT),
    (dict(dt=4e-4), _lib.KG_ERR_UNSTABLE),          # dt/dx^2 = 0.44 > 3/8
    (dict(eps_bar=0.1, tau=1e-5), _lib.KG_ERR_UNSTABLE),
])
def test_validate_rejects(L, bad, code):
    with pytest.raises(_lib.KGError) as e:
        kb.validate(64, 64, kb.default_params(**bad))
    assert e.value.code == code
    assert L.kg_last_error()


def test_validate_accepts_configs(L):
    kb.validate(300, 300, kb.default_params(delta=0.02, aniso=4, theta0=0.2, noise_a=0.01, boundary=1))
    kb.validate(4096, 4096, kb.default_params(delta=0.04, aniso=6, K=2.0, theta0=0.2, noise_a=0.01, boundary=1))
    with pytest.raises(_lib.KGError):
        kb.validate(2, 64, kb.default_params())


def test_reference_range_mask(L):
    m = ctypes.c_uint32(0)
    p = kb.default_params()
    _lib.check(L.kg_reference_range_mask(ctypes.byref(p), ctypes.byref(m)))
    assert m.value == 0
    p = p.replace(K=2.0, delta=0.5)
    _lib.check(L.kg_reference_range_mask(ctypes.byref(p), ctypes.byref(m)))
    assert m.value == (1 << 3) | (1 << 4)


def test_product_philox_matches_published_vectors(L):
    path = os.path.join(os.path.dirname(__file__), "golden", "philox4x32_10_kat.txt")
    rows = [l.split() for l in open(path) if l.strip() and not l.startswith("#")]
    for r in rows:
        v = [int(x, 16) for x in r]
        assert kb.philox4x32_10(v[0:4], v[4:6]).tolist() == v[6:10]


def test_null_handle_errors(L):
    assert L.kg_step(None, 1) == _lib.KG_ERR_INVALID_ARGUMENT
    assert L.kg_destroy(None) == _lib.KG_ERR_INVALID_ARGUMENT
    assert b"null" in L.kg_last_error()
    u, v = ctypes.c_uint64(), ctypes.c_uint64()
    assert L.kg_get_skip_stats(None, ctypes.byref(u), ctypes.byref(v)) == _lib.KG_ERR_INVALID_ARGUMENT


def test_create_without_gpu_fails_loudly(L):
    if kb.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(_lib.KGError) as e:
        kb.Solver(32, 32)
    assert e.value.code == _lib.KG_ERR_DEVICE
