"""Stream capture of kg_step (needs a GPU).

Every step launch carries its step count (the noise counter) and the
buffer parity in its arguments, so a captured graph would replay the same
step: kg_step refuses while its stream is being captured, and the handle
steps normally once the capture has ended.  (Batches in KG_BATCH_GRAPH mode
read both from device memory and replay correctly: tests/test_gpu_batch.py.)"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def kb():
    import crystalgrowth_amd as kb
    if kb.device_count() < 1:
        pytest.fail("no GPU visible")
    return kb


def test_step_refuses_stream_capture(kb):
    from crystalgrowth_amd._lib import KGError, lib
    # the HIP runtime the library itself uses (symbol lookup through its
    # handle searches its dependencies): a process may hold another copy
    hip = lib()
    hip.hipStreamBeginCapture.argtypes = [ctypes.c_void_p, ctypes.c_int]
    hip.hipStreamEndCapture.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]
    hip.hipGraphDestroy.argtypes = [ctypes.c_void_p]
    p = kb.default_params(noise_a=0.01, seed=5)
    with kb.Solver(96, 80, "f32", p) as s, kb.Solver(96, 80, "f32", p) as ref:
        s.reset_disk(48, 40, 5.0)
        ref.reset_disk(48, 40, 5.0)
        s.step(3)
        ref.step(3)
        s.synchronize()
        st = ctypes.c_void_p(s.stream())
        assert hip.hipStreamBeginCapture(st, 2) == 0  # hipStreamCaptureModeRelaxed
        graph = ctypes.c_void_p()
        try:
            with pytest.raises(KGError, match="captured"):
                s.step(1)
        finally:
            rc = hip.hipStreamEndCapture(st, ctypes.byref(graph))
            if graph.value:
                hip.hipGraphDestroy(graph)
        assert rc == 0
        assert s.step_count == 3
        s.step(5)
        ref.step(5)
        for name, x, y in zip(("phi", "T", "theta"), s.fields(), ref.fields()):
            assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), name
