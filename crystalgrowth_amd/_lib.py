"""ctypes binding of libkobayashi_gpu.so (include/kobayashi_gpu.h).

The library is the product: there is no CPU fallback.  If it is missing the
import of the binding fails with a message that names the build command.
"""
from __future__ import annotations

import ctypes
import os
import re

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libkobayashi_gpu.so")
HEADER = os.path.join(os.path.dirname(PKG), "include", "kobayashi_gpu.h")

KG_OK = 0
KG_ERR_INVALID_ARGUMENT = -1
KG_ERR_UNSTABLE = -2
KG_ERR_OUT_OF_MEMORY = -3
KG_ERR_DEVICE = -4
KG_ERR_SIZE = -5
KG_F32, KG_F64 = 0, 1
KG_PERIODIC, KG_ZERO_FLUX = 0, 1
KG_KERNEL_FUSED, KG_KERNEL_TWO_PASS, KG_KERNEL_FUSED_DENSE = 0, 1, 2
KG_BATCH_LAUNCH, KG_BATCH_GRAPH, KG_BATCH_GRAPH_LISTS = 0, 1, 2


class KGError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class Params(ctypes.Structure):
    """kg_params (include/kobayashi_gpu.h)."""

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
            if not hasattr(q, k):
                raise AttributeError(f"kg_params has no field {k!r}")
            setattr(q, k, v)
        return q

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class Grads(ctypes.Structure):
    """kg_grads (include/kobayashi_gpu.h): loss and dL/d(parameter)."""

    _fields_ = [(n, ctypes.c_double) for n in (
        "loss", "d_tau", "d_eps_bar", "d_K", "d_delta", "d_theta0", "d_alpha", "d_gamma",
        "d_teq", "d_noise_a")]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class Stats(ctypes.Structure):
    _fields_ = [
        ("sum_phi", ctypes.c_double), ("sum_T", ctypes.c_double),
        ("sum_T_minus_K_phi", ctypes.c_double), ("min_phi", ctypes.c_double),
        ("max_phi", ctypes.c_double), ("solid_fraction", ctypes.c_double),
        ("nonfinite", ctypes.c_int64), ("cells", ctypes.c_int64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


H = ctypes.c_void_p
P = ctypes.POINTER(Params)
I = ctypes.c_int
PI = ctypes.POINTER(ctypes.c_int)
VP = ctypes.c_void_p
SZ = ctypes.c_size_t
U64 = ctypes.c_uint64
U32P = ctypes.POINTER(ctypes.c_uint32)

# every entry point of include/kobayashi_gpu.h with its argument types
SIGNATURES = {
    "kg_abi_version": [],
    "kg_last_error": [],
    "kg_device_count": [PI],
    "kg_default_params": [P],
    "kg_validate": [I, I, P],
    "kg_reference_range_mask": [P, U32P],
    "kg_philox4x32_10": [U32P, U32P, U32P],
    "kg_create": [I, I, I, P, ctypes.POINTER(H)],
    "kg_destroy": [H],
    "kg_reset": [H],
    "kg_reset_disk": [H, I, I, ctypes.c_double],
    "kg_reset_seeds": [H, I, ctypes.c_double, I, U64, PI],
    "kg_set_params": [H, P],
    "kg_get_params": [H, P],
    "kg_get_size": [H, PI, PI, PI],
    "kg_step": [H, I],
    "kg_step_timed": [H, I, ctypes.POINTER(ctypes.c_double)],
    "kg_synchronize": [H],
    "kg_get_step_count": [H, ctypes.POINTER(U64)],
    "kg_get_skip_stats": [H, ctypes.POINTER(U64), ctypes.POINTER(U64)],
    "kg_set_step_count": [H, U64],
    "kg_set_kernel": [H, I],
    "kg_get_phase": [H, VP, SZ],
    "kg_get_temperature": [H, VP, SZ],
    "kg_get_angle": [H, VP, SZ],
    "kg_set_fields": [H, VP, VP, VP, SZ],
    "kg_device_fields": [H, ctypes.POINTER(VP), ctypes.POINTER(VP), ctypes.POINTER(VP)],
    "kg_get_stream": [H, ctypes.POINTER(VP)],
    "kg_reduce": [H, ctypes.POINTER(Stats)],
    "kg_colormap": [H, VP, SZ, I],
    "kg_colormap_rgba8": [H, VP, SZ, I],
    "kg_colormap_rgba8_noalloc": [H, VP, SZ, I],
    "kg_front_profile": [H, ctypes.c_double, ctypes.c_double, I, ctypes.c_double, VP],
    "kg_adjoint": [H, I, I, VP, SZ, ctypes.POINTER(Grads), VP, VP],
    "kg_batch_create": [I, PI, PI, I, P, ctypes.POINTER(H)],
    "kg_batch_destroy": [H],
    "kg_batch_count": [H, PI],
    "kg_batch_set_params": [H, I, P],
    "kg_batch_get_params": [H, I, P],
    "kg_batch_adjoint": [H, I, I, ctypes.POINTER(VP), ctypes.POINTER(Grads), ctypes.POINTER(VP)],
    "kg_batch_reset_disk": [H, I, I, I, ctypes.c_double],
    "kg_batch_reset_seeds": [H, I, I, ctypes.c_double, I, U64],
    "kg_batch_set_fields": [H, I, VP, VP, VP, SZ],
    "kg_batch_get_phase": [H, I, VP, SZ],
    "kg_batch_get_temperature": [H, I, VP, SZ],
    "kg_batch_get_angle": [H, I, VP, SZ],
    "kg_batch_reduce": [H, I, ctypes.POINTER(Stats)],
    "kg_batch_front_profile": [H, I, ctypes.c_double, ctypes.c_double, I, ctypes.c_double, VP],
    "kg_batch_set_mode": [H, I, I],
    "kg_batch_get_mode": [H, PI, PI],
    "kg_batch_set_chunk_shape": [H, I, I],
    "kg_batch_get_chunk_shape": [H, PI, PI],
    "kg_batch_autotune": [H, I, PI, PI],
    "kg_batch_step": [H, I],
    "kg_batch_step_timed": [H, I, ctypes.POINTER(ctypes.c_double)],
    "kg_batch_synchronize": [H],
    "kg_batch_get_step_count": [H, ctypes.POINTER(U64)],
    "kg_batch_get_skip_stats": [H, ctypes.POINTER(U64), ctypes.POINTER(U64)],
}


def header_symbols(path: str = HEADER) -> list[str]:
    """Names of the functions the public header declares."""
    text = open(path).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(kg_\w+)\s*\(", text, re.M)))


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m crystalgrowth_amd.build` "
            "(hipcc for gfx950); there is no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_char_p if name == "kg_last_error" else ctypes.c_int
    _lib = L
    return L


def check(code: int) -> None:
    if code != KG_OK:
        raise KGError(code, lib().kg_last_error().decode(errors="replace"))
