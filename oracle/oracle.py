"""ctypes binding of the CPU oracle (oracle/kobayashi_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product package
(crystalgrowth_amd) never imports this module.

The oracle restates src/Kobayashi.cpp:125-221 (one explicit-Euler step of the
Kobayashi model) in float and double; see kobayashi_oracle.h for the mapping.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# KGO_ORACLE_LIB selects another build of the same source (the mutant builds
# of tests/test_oracle_mutants.py); the default is the plain build.
LIB_PATH = os.environ.get("KGO_ORACLE_LIB") or os.path.join(HERE, "build", "libkg_oracle.so")

PERIODIC = 0
ZERO_FLUX = 1


class Params(ctypes.Structure):
    """Mirror of kgo_params (kobayashi_oracle.h)."""

    _fields_ = [
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dt", ctypes.c_double),
        ("tau", ctypes.c_double), ("eps_bar", ctypes.c_double), ("mu", ctypes.c_double),
        ("K", ctypes.c_double), ("delta", ctypes.c_double), ("aniso", ctypes.c_double),
        ("alpha", ctypes.c_double), ("gamma", ctypes.c_double), ("teq", ctypes.c_double),
        ("theta0", ctypes.c_double), ("noise_a", ctypes.c_double),
        ("boundary", ctypes.c_int32), ("instance", ctypes.c_uint32),
        ("seed", ctypes.c_uint64),
    ]

    def replace(self, **kw) -> "Params":
        q = Params()
        ctypes.pointer(q)[0] = self
        for k, v in kw.items():
            setattr(q, k, v)
        return q

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


def build() -> None:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None
_lib_omp = None
OMP_PATH = os.path.join(HERE, "build", "libkg_oracle_omp.so")


def lib_parallel():
    """The same oracle with its row loops on OpenMP threads (bitwise the
    serial results); used for the all-cores CPU timing."""
    global _lib_omp
    if _lib_omp is None:
        if not os.path.exists(OMP_PATH):
            build()
        L = ctypes.CDLL(OMP_PATH)
        P = ctypes.POINTER(Params)
        for suf, dt in (("f32", np.float32), ("f64", np.float64)):
            ap = np.ctypeslib.ndpointer(dt, flags="C_CONTIGUOUS")
            getattr(L, f"kgo_run_{suf}").argtypes = [
                ctypes.c_int, ctypes.c_int, P, ap, ap, ap, ctypes.c_uint32, ctypes.c_int]
        L.kgo_threads.restype = ctypes.c_int
        _lib_omp = L
    return _lib_omp


def run_parallel(state: "State", p: Params, nsteps: int, step0: int = 0) -> "State":
    s = state.copy()
    ny, nx = s.shape
    suf = _dtype_suffix(s.phi.dtype)
    getattr(lib_parallel(), f"kgo_run_{suf}")(nx, ny, ctypes.byref(p), s.phi, s.T, s.theta,
                                              step0, nsteps)
    return s


def threads_parallel() -> int:
    return int(lib_parallel().kgo_threads())


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER(Params)
        f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
        f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
        i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        L.kgo_default_params.argtypes = [P]
        L.kgo_philox4x32_10.argtypes = [u32p, u32p, u32p]
        L.kgo_noise.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_uint32]
        L.kgo_noise.restype = ctypes.c_double
        for suf, ap in (("f32", f32p), ("f64", f64p)):
            getattr(L, f"kgo_step_{suf}").argtypes = [
                ctypes.c_int, ctypes.c_int, P, ap, ap, ap, ctypes.c_uint32, ctypes.c_void_p]
            getattr(L, f"kgo_run_{suf}").argtypes = [
                ctypes.c_int, ctypes.c_int, P, ap, ap, ap, ctypes.c_uint32, ctypes.c_int]
            getattr(L, f"kgo_pass1_{suf}").argtypes = [
                ctypes.c_int, ctypes.c_int, P, ap, ap, ap, ap]
        L.kgo_init_plus_f64.argtypes = [ctypes.c_int, ctypes.c_int, f64p, f64p, f64p]
        L.kgo_init_disk_f64.argtypes = [ctypes.c_int, ctypes.c_int, f64p, f64p, f64p,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.kgo_init_seeds_f64.argtypes = [ctypes.c_int, ctypes.c_int, f64p, f64p, f64p,
                                         ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                         ctypes.c_uint64, ctypes.c_void_p]
        _lib = L
    return _lib


def default_params(**overrides) -> Params:
    p = Params()
    lib().kgo_default_params(ctypes.byref(p))
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


def philox4x32_10(ctr, key) -> np.ndarray:
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, np.uint32)
    lib().kgo_philox4x32_10(c, k, out)
    return out


def noise(p: Params, i: int, j: int, step: int) -> float:
    return lib().kgo_noise(ctypes.byref(p), i, j, step)


def _dtype_suffix(dtype) -> str:
    dt = np.dtype(dtype)
    if dt == np.float32:
        return "f32"
    if dt == np.float64:
        return "f64"
    raise TypeError(f"oracle supports float32/float64, got {dt}")


class State:
    """(phi, T, theta) arrays of shape (ny, nx), x fastest."""

    def __init__(self, phi, T, theta):
        self.phi, self.T, self.theta = phi, T, theta

    @property
    def shape(self):
        return self.phi.shape

    def astype(self, dtype) -> "State":
        return State(*(np.ascontiguousarray(a, dtype=dtype) for a in (self.phi, self.T, self.theta)))

    def copy(self) -> "State":
        return State(self.phi.copy(), self.T.copy(), self.theta.copy())


def zeros(nx: int, ny: int, dtype=np.float64) -> State:
    return State(*(np.zeros((ny, nx), dtype) for _ in range(3)))


def init_plus(nx: int, ny: int, dtype=np.float64) -> State:
    s = zeros(nx, ny)
    lib().kgo_init_plus_f64(nx, ny, s.phi, s.T, s.theta)
    return s.astype(dtype)


def init_disk(nx: int, ny: int, cx: int, cy: int, r: float, dtype=np.float64) -> State:
    s = zeros(nx, ny)
    lib().kgo_init_disk_f64(nx, ny, s.phi, s.T, s.theta, cx, cy, r)
    return s.astype(dtype)


def init_seeds(nx: int, ny: int, count: int, r: float, margin: int, seed: int,
               dtype=np.float64):
    s = zeros(nx, ny)
    cxy = np.zeros(2 * count, np.int32)
    lib().kgo_init_seeds_f64(nx, ny, s.phi, s.T, s.theta, count, r, margin, seed,
                             cxy.ctypes.data_as(ctypes.c_void_p))
    return s.astype(dtype), cxy.reshape(count, 2)


def step(state: State, p: Params, step_index: int = 0) -> State:
    """One step in the state's dtype; returns a new State."""
    s = state.copy()
    ny, nx = s.shape
    suf = _dtype_suffix(s.phi.dtype)
    getattr(lib(), f"kgo_step_{suf}")(nx, ny, ctypes.byref(p), s.phi, s.T, s.theta,
                                      step_index, None)
    return s


def run(state: State, p: Params, nsteps: int, step0: int = 0) -> State:
    s = state.copy()
    ny, nx = s.shape
    suf = _dtype_suffix(s.phi.dtype)
    getattr(lib(), f"kgo_run_{suf}")(nx, ny, ctypes.byref(p), s.phi, s.T, s.theta,
                                     step0, nsteps)
    return s


def pass1(state: State, p: Params) -> dict:
    """Pass-1 arrays of one step (state untouched)."""
    ny, nx = state.shape
    suf = _dtype_suffix(state.phi.dtype)
    out = np.zeros((7, ny, nx), state.phi.dtype)
    getattr(lib(), f"kgo_pass1_{suf}")(nx, ny, ctypes.byref(p), state.phi, state.T,
                                       state.theta, out)
    names = ["gx", "gy", "lap_phi", "lap_T", "theta", "eps", "eps_deriv"]
    return {n: out[k] for k, n in enumerate(names)}
