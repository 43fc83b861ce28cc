"""The headless C++ driver (csrc/kobayashi_headless.cpp: the reference's
WinMain loop without a window, on the C++ Kobayashi class over the C ABI)
runs on the GPU and ends where the Python class over the same ABI ends."""
import json
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "crystalgrowth_amd", "bin", "kobayashi_headless")


def test_headless_driver_matches_python_class(tmp_path):
    import crystalgrowth_amd as kb
    if kb.device_count() < 1:
        pytest.fail("no GPU visible")
    if not os.path.exists(BIN):
        pytest.fail("headless driver not built (crystalgrowth_amd/build.py)")
    ppm = tmp_path / "frame.ppm"
    r = subprocess.run([BIN, "--size", "64", "--frames", "20", "--ppm", str(ppm)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["size"] == 64 and out["frames"] == 20 and out["steps"] == 200
    sim = kb.Kobayashi(64, 64, 1e-4)
    try:
        for _ in range(20):
            sim.iUpdate()
        st = sim.stats()
    finally:
        sim.close()
    assert out["sum_phi"] == st["sum_phi"]
    assert out["sum_T_minus_K_phi"] == st["sum_T_minus_K_phi"]
    assert out["nonfinite"] == 0
    data = ppm.read_bytes()
    header = b"P6\n64 64\n255\n"
    assert data.startswith(header) and len(data) == len(header) + 64 * 64 * 3
    px = np.frombuffer(data[len(header):], np.uint8)
    assert px.max() > px.min()   # the nucleus shows
