"""Python host API over the C ABI.

`Solver` is the general handle (any size, FP32/FP64, periodic or zero-flux,
noise, seeded initial states).  `Kobayashi` keeps the reference class's
simulation surface (src/Kobayashi.h:35-59): `Kobayashi(x, y, timeStep)`,
`iUpdate()` = 10 steps (src/Kobayashi.cpp:227-239), `iResetSimulationState()`,
`iGetObjectCount()`, `iIsUpdated()`, `iUpdateConstantBuffer()` colours
(:309-345), plus the phase/temperature/angle accessors the headless use needs.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import (KG_BATCH_GRAPH, KG_BATCH_GRAPH_LISTS, KG_BATCH_LAUNCH, KG_F32, KG_F64,
                   KG_KERNEL_FUSED, KG_KERNEL_FUSED_DENSE, KG_KERNEL_TWO_PASS, Grads, Params, Stats, check, lib)

_PREC = {"f32": KG_F32, "float32": KG_F32, "f64": KG_F64, "float64": KG_F64}
_KERNELS = {"fused": KG_KERNEL_FUSED, "two_pass": KG_KERNEL_TWO_PASS,
            "fused_dense": KG_KERNEL_FUSED_DENSE}


def default_params(**overrides) -> Params:
    """Reference defaults (src/Kobayashi.cpp:61-63, :76-84, dt of src/main.cpp:16)."""
    p = Params()
    check(lib().kg_default_params(ctypes.byref(p)))
    return p.replace(**overrides) if overrides else p


def validate(nx: int, ny: int, params: Params) -> None:
    check(lib().kg_validate(nx, ny, ctypes.byref(params)))


def device_count() -> int:
    n = ctypes.c_int(0)
    check(lib().kg_device_count(ctypes.byref(n)))
    return n.value


def philox4x32_10(ctr, key) -> np.ndarray:
    c = (ctypes.c_uint32 * 4)(*[int(x) for x in ctr])
    k = (ctypes.c_uint32 * 2)(*[int(x) for x in key])
    o = (ctypes.c_uint32 * 4)()
    check(lib().kg_philox4x32_10(c, k, o))
    return np.array(list(o), dtype=np.uint32)


class Solver:
    """One device-resident grid (phi, T, theta) and its parameters."""

    def __init__(self, nx: int, ny: int, precision: str = "f32", params: Params | None = None,
                 **overrides):
        if precision not in _PREC:
            raise ValueError(f"precision must be one of {sorted(_PREC)}")
        p = params if params is not None else default_params()
        if overrides:
            p = p.replace(**overrides)
        self.nx, self.ny = int(nx), int(ny)
        self.precision = _PREC[precision]
        self.dtype = np.float64 if self.precision == KG_F64 else np.float32
        self._h = ctypes.c_void_p()
        check(lib().kg_create(self.nx, self.ny, self.precision, ctypes.byref(p),
                              ctypes.byref(self._h)))

    # -- lifecycle --------------------------------------------------------
    def close(self) -> None:
        if self._h:
            check(lib().kg_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- parameters -------------------------------------------------------
    @property
    def params(self) -> Params:
        p = Params()
        check(lib().kg_get_params(self._h, ctypes.byref(p)))
        return p

    @params.setter
    def params(self, p: Params) -> None:
        check(lib().kg_set_params(self._h, ctypes.byref(p)))

    def set_params(self, **kw) -> None:
        self.params = self.params.replace(**kw)

    def set_kernel(self, name: str) -> None:
        check(lib().kg_set_kernel(self._h, _KERNELS[name]))

    # -- initial states ---------------------------------------------------
    def reset(self) -> None:
        check(lib().kg_reset(self._h))

    def reset_disk(self, cx: int, cy: int, r: float) -> None:
        check(lib().kg_reset_disk(self._h, int(cx), int(cy), float(r)))

    def reset_seeds(self, count: int, r: float, margin: int, seed: int) -> np.ndarray:
        cxy = (ctypes.c_int * (2 * count))()
        check(lib().kg_reset_seeds(self._h, int(count), float(r), int(margin), int(seed), cxy))
        return np.array(list(cxy), dtype=np.int32).reshape(count, 2)

    # -- stepping ---------------------------------------------------------
    def step(self, n: int = 1) -> None:
        check(lib().kg_step(self._h, int(n)))

    def step_timed(self, n: int) -> float:
        ms = ctypes.c_double(0.0)
        check(lib().kg_step_timed(self._h, int(n), ctypes.byref(ms)))
        return ms.value

    def synchronize(self) -> None:
        check(lib().kg_synchronize(self._h))

    @property
    def step_count(self) -> int:
        v = ctypes.c_uint64(0)
        check(lib().kg_get_step_count(self._h, ctypes.byref(v)))
        return v.value

    @step_count.setter
    def step_count(self, v: int) -> None:
        check(lib().kg_set_step_count(self._h, int(v)))

    def skip_stats(self) -> dict:
        """Launches that covered only the live chunks, and liveness answers
        dropped as not yet visible (kg_get_skip_stats)."""
        listed, stale = ctypes.c_uint64(0), ctypes.c_uint64(0)
        check(lib().kg_get_skip_stats(self._h, ctypes.byref(listed), ctypes.byref(stale)))
        return {"listed_launches": listed.value, "stale_answers": stale.value}

    # -- fields -----------------------------------------------------------
    def _get(self, fn) -> np.ndarray:
        a = np.empty((self.ny, self.nx), self.dtype)
        check(fn(self._h, a.ctypes.data_as(ctypes.c_void_p), a.nbytes))
        return a

    def phase(self) -> np.ndarray:
        return self._get(lib().kg_get_phase)

    def temperature(self) -> np.ndarray:
        return self._get(lib().kg_get_temperature)

    def angle(self) -> np.ndarray:
        return self._get(lib().kg_get_angle)

    def fields(self):
        return self.phase(), self.temperature(), self.angle()

    def set_fields(self, phi, T, theta=None) -> None:
        shape = (self.ny, self.nx)
        arrs = [np.ascontiguousarray(a, self.dtype) for a in (phi, T)]
        th = None if theta is None else np.ascontiguousarray(theta, self.dtype)
        for a in arrs + ([th] if th is not None else []):
            if a.shape != shape:
                raise ValueError(f"field shape {a.shape} != {shape}")
        check(lib().kg_set_fields(self._h, arrs[0].ctypes.data_as(ctypes.c_void_p),
                                  arrs[1].ctypes.data_as(ctypes.c_void_p),
                                  None if th is None else th.ctypes.data_as(ctypes.c_void_p),
                                  arrs[0].nbytes))

    def device_fields(self):
        """(phi, T, theta) device addresses, valid until the next step/reset."""
        p, t, th = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        check(lib().kg_device_fields(self._h, ctypes.byref(p), ctypes.byref(t), ctypes.byref(th)))
        return p.value, t.value, th.value

    def stream(self) -> int:
        s = ctypes.c_void_p()
        check(lib().kg_get_stream(self._h, ctypes.byref(s)))
        return s.value or 0

    # -- reductions / readback -------------------------------------------
    def stats(self) -> dict:
        s = Stats()
        check(lib().kg_reduce(self._h, ctypes.byref(s)))
        return s.as_dict()

    def front_profile(self, cx: float, cy: float, n: int = 720, level: float = 0.5) -> np.ndarray:
        """Distance from (cx, cy) to the phi = level front along n rays, on the
        device (the algorithm of morphology.radius_profile)."""
        out = np.empty(int(n), np.float64)
        check(lib().kg_front_profile(self._h, float(cx), float(cy), int(n), float(level),
                                     out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def colormap(self, order: int = 0) -> np.ndarray:
        out = np.empty((self.nx * self.ny, 4), np.float32)
        check(lib().kg_colormap(self._h, out.ctypes.data_as(ctypes.c_void_p), out.nbytes, order))
        return out

    def colormap_rgba8(self, order: int = 0) -> np.ndarray:
        """The colours as RGBA8 (kg_colormap_rgba8), shape (cells, 4) uint8."""
        out = np.empty((self.nx * self.ny, 4), np.uint8)
        check(lib().kg_colormap_rgba8(self._h, out.ctypes.data_as(ctypes.c_void_p), out.nbytes,
                                      order))
        return out

    def colormap_rgba8_noalloc(self, order: int = 0, out: np.ndarray | None = None) -> np.ndarray:
        """kg_colormap_rgba8 without device allocations (kg_colormap_rgba8_noalloc):
        grids up to KG_COLORMAP_NOALLOC_MAX_CELLS cells; `out` may be reused."""
        if out is None:
            out = np.empty((self.nx * self.ny, 4), np.uint8)
        if out.dtype != np.uint8 or out.nbytes != self.nx * self.ny * 4 or \
                not out.flags["C_CONTIGUOUS"]:
            raise ValueError("out must be a C-contiguous uint8 array of 4 bytes per cell")
        check(lib().kg_colormap_rgba8_noalloc(self._h, out.ctypes.data_as(ctypes.c_void_p),
                                              out.nbytes, order))
        return out

    # -- adjoint ----------------------------------------------------------
    def adjoint(self, nsteps: int, target, checkpoint_every: int = 50, want_T0: bool = False,
                want_phi0: bool = False):
        """Loss sum (phi_N - target)^2 and its gradients from the current state
        (kg_adjoint).  Returns (grads dict, dL/dT0 or None, dL/dphi0 or None);
        the solver is left at the final forward state."""
        tgt = np.ascontiguousarray(target, self.dtype)
        if tgt.shape != (self.ny, self.nx):
            raise ValueError("target shape must be (ny, nx)")
        g = Grads()
        dT = np.empty((self.ny, self.nx), self.dtype) if want_T0 else None
        dP = np.empty((self.ny, self.nx), self.dtype) if want_phi0 else None
        check(lib().kg_adjoint(self._h, int(nsteps), int(checkpoint_every),
                               tgt.ctypes.data_as(ctypes.c_void_p), tgt.nbytes, ctypes.byref(g),
                               None if dT is None else dT.ctypes.data_as(ctypes.c_void_p),
                               None if dP is None else dP.ctypes.data_as(ctypes.c_void_p)))
        return g.as_dict(), dT, dP


class Kobayashi(Solver):
    """The reference class's simulation surface (src/Kobayashi.h:32-67) without
    Win32/D3D: the GUI's 10 steps per frame, reset, object count and colours."""

    STEPS_PER_UPDATE = 10  # src/Kobayashi.cpp:230

    def __init__(self, x: int, y: int, timeStep: float, precision: str = "f32"):
        super().__init__(x, y, precision, default_params(dt=float(np.float32(timeStep))))
        self._updateFlag = True
        self.simFrame = 0

    def iUpdate(self) -> None:
        self.step(self.STEPS_PER_UPDATE)
        self.simFrame += 1

    def iResetSimulationState(self) -> None:
        self.reset()
        self.simFrame = 0

    def iGetObjectCount(self):
        return (self.nx, self.ny, 0)

    def iIsUpdated(self) -> bool:
        return self._updateFlag

    def iUpdateConstantBuffer(self) -> np.ndarray:
        """Colours of every constant buffer entry, in the reference's order."""
        return self.colormap(order=1)


_MODE_NAMES = {KG_BATCH_LAUNCH: "launch", KG_BATCH_GRAPH: "graph",
               KG_BATCH_GRAPH_LISTS: "graph_lists"}


class Batch:
    """An ensemble of independent grids of any sizes and parameters stepped
    together (kg_batch_* of include/kobayashi_gpu.h).  Member m evolves bit for
    bit like a Solver with the same size, precision, parameters and state."""

    def __init__(self, sizes, params, precision: str = "f32"):
        sizes = [(int(nx), int(ny)) for nx, ny in sizes]
        if isinstance(params, Params):
            params = [params] * len(sizes)
        if len(params) != len(sizes):
            raise ValueError("one kg_params per member")
        n = len(sizes)
        self.sizes = sizes
        self.precision = _PREC[precision]
        self.dtype = np.float64 if self.precision == KG_F64 else np.float32
        nx = (ctypes.c_int * n)(*[s[0] for s in sizes])
        ny = (ctypes.c_int * n)(*[s[1] for s in sizes])
        arr = (Params * n)(*params)
        self._h = ctypes.c_void_p()
        check(lib().kg_batch_create(n, nx, ny, self.precision, arr, ctypes.byref(self._h)))

    def __len__(self):
        return len(self.sizes)

    def close(self) -> None:
        if self._h:
            check(lib().kg_batch_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def reset_disk(self, member: int = -1, cx: int = -1, cy: int = -1, r: float = 5.0) -> None:
        check(lib().kg_batch_reset_disk(self._h, member, cx, cy, float(r)))

    def reset_seeds(self, count: int, r: float, margin: int, seed: int, member: int = -1) -> None:
        check(lib().kg_batch_reset_seeds(self._h, member, count, float(r), margin, int(seed)))

    def set_fields(self, member: int, phi, T, theta=None) -> None:
        nx, ny = self.sizes[member]
        a = [np.ascontiguousarray(x, self.dtype).reshape(ny, nx) for x in (phi, T)]
        th = None if theta is None else np.ascontiguousarray(theta, self.dtype).reshape(ny, nx)
        check(lib().kg_batch_set_fields(self._h, member, a[0].ctypes.data_as(ctypes.c_void_p),
                                        a[1].ctypes.data_as(ctypes.c_void_p),
                                        None if th is None else th.ctypes.data_as(ctypes.c_void_p),
                                        a[0].nbytes))

    def _get(self, fn, member):
        nx, ny = self.sizes[member]
        a = np.empty((ny, nx), self.dtype)
        check(fn(self._h, member, a.ctypes.data_as(ctypes.c_void_p), a.nbytes))
        return a

    def phase(self, member: int) -> np.ndarray:
        return self._get(lib().kg_batch_get_phase, member)

    def temperature(self, member: int) -> np.ndarray:
        return self._get(lib().kg_batch_get_temperature, member)

    def angle(self, member: int) -> np.ndarray:
        return self._get(lib().kg_batch_get_angle, member)

    def fields(self, member: int):
        return self.phase(member), self.temperature(member), self.angle(member)

    def stats(self, member: int) -> dict:
        s = Stats()
        check(lib().kg_batch_reduce(self._h, member, ctypes.byref(s)))
        return s.as_dict()

    def front_profile(self, member: int, cx: float, cy: float, n: int = 720,
                      level: float = 0.5) -> np.ndarray:
        out = np.empty(int(n), np.float64)
        check(lib().kg_batch_front_profile(self._h, int(member), float(cx), float(cy), int(n),
                                           float(level), out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def set_params(self, member: int, p: Params) -> None:
        check(lib().kg_batch_set_params(self._h, member, ctypes.byref(p)))

    def params(self, member: int) -> Params:
        p = Params()
        check(lib().kg_batch_get_params(self._h, member, ctypes.byref(p)))
        return p

    def adjoint(self, nsteps: int, targets, checkpoint_every: int = 50, want_T0: bool = False):
        """kg_batch_adjoint: per-member (grads dict, dL/dT0 or None) from the
        current state; the batch is left at the final forward state."""
        n = len(self.sizes)
        tg = []
        for m, t in enumerate(targets):
            nx, ny = self.sizes[m]
            tg.append(np.ascontiguousarray(t, self.dtype).reshape(ny, nx))
        tptr = (ctypes.c_void_p * n)(*[t.ctypes.data for t in tg])
        dT = [np.empty((ny, nx), self.dtype) for nx, ny in self.sizes] if want_T0 else None
        dptr = (ctypes.c_void_p * n)(*[d.ctypes.data for d in dT]) if want_T0 else None
        out = (Grads * n)()
        check(lib().kg_batch_adjoint(self._h, int(nsteps), int(checkpoint_every), tptr, out, dptr))
        return [g.as_dict() for g in out], dT

    def set_mode(self, mode: str = "launch", graph_steps: int = 0) -> None:
        m = {"launch": KG_BATCH_LAUNCH, "graph": KG_BATCH_GRAPH,
             "graph_lists": KG_BATCH_GRAPH_LISTS}[mode]
        check(lib().kg_batch_set_mode(self._h, m, int(graph_steps)))

    def mode(self):
        m, g = ctypes.c_int(), ctypes.c_int()
        check(lib().kg_batch_get_mode(self._h, ctypes.byref(m), ctypes.byref(g)))
        return _MODE_NAMES[m.value], g.value

    CHUNK_SHAPES = ((2, 2), (2, 4), (4, 2), (4, 1), (8, 1))

    def set_chunk_shape(self, width: int, height: int) -> None:
        """Cut every member into chunks of width x height tiles (one of
        CHUNK_SHAPES; kg_batch_set_chunk_shape)."""
        check(lib().kg_batch_set_chunk_shape(self._h, int(width), int(height)))

    @property
    def chunk_shape(self):
        w, h = ctypes.c_int(), ctypes.c_int()
        check(lib().kg_batch_get_chunk_shape(self._h, ctypes.byref(w), ctypes.byref(h)))
        return w.value, h.value

    def autotune(self, trial_steps: int = 16):
        """Times every launch form in every chunk shape from the current
        state, keeps the fastest and restores the state; returns the launch
        form (chunk_shape gives the shape)."""
        m, g = ctypes.c_int(), ctypes.c_int()
        check(lib().kg_batch_autotune(self._h, int(trial_steps), ctypes.byref(m), ctypes.byref(g)))
        return _MODE_NAMES[m.value], g.value

    def step(self, n: int = 1) -> None:
        check(lib().kg_batch_step(self._h, int(n)))

    def step_timed(self, n: int) -> float:
        ms = ctypes.c_double()
        check(lib().kg_batch_step_timed(self._h, int(n), ctypes.byref(ms)))
        return ms.value

    def synchronize(self) -> None:
        check(lib().kg_batch_synchronize(self._h))

    def skip_stats(self) -> dict:
        """Launches that covered only the live chunks, and liveness answers
        dropped as not yet visible (kg_batch_get_skip_stats)."""
        listed, stale = ctypes.c_uint64(0), ctypes.c_uint64(0)
        check(lib().kg_batch_get_skip_stats(self._h, ctypes.byref(listed), ctypes.byref(stale)))
        return {"listed_launches": listed.value, "stale_answers": stale.value}

    @property
    def step_count(self) -> int:
        v = ctypes.c_uint64()
        check(lib().kg_batch_get_step_count(self._h, ctypes.byref(v)))
        return v.value
