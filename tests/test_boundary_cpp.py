"""The C++ class keeps the reference's ISimulation signatures (no GPU needed:
compiled, not run).

src/Kobayashi.h:35-62 declares the overrides the viewer calls; a program
that takes the address of each member as a pointer of the reference's exact
type compiles only if the headless class has those signatures."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "crystalgrowth_amd", "csrc")
INCLUDE = os.path.join(ROOT, "include")

PROGRAM = r"""
#include <type_traits>
#include "Kobayashi.h"
using CB = std::vector<ConstantBuffer>;
// src/Kobayashi.h:35-62, with UINT = unsigned int
static_assert(std::is_base_of<ISimulation, Kobayashi>::value, "ISimulation subclass");
static_assert(std::is_constructible<Kobayashi, int, int, float>::value, "Kobayashi(int, int, float)");
void (Kobayashi::*update)() = &Kobayashi::iUpdate;
void (Kobayashi::*reset)(CB &) = &Kobayashi::iResetSimulationState;
std::vector<Vertex> &(Kobayashi::*vertices)() = &Kobayashi::iGetVertices;
std::vector<unsigned int> &(Kobayashi::*indices)() = &Kobayashi::iGetIndices;
unsigned int (Kobayashi::*vbsize)() = &Kobayashi::iGetVertexBufferSize;
unsigned int (Kobayashi::*ibsize)() = &Kobayashi::iGetIndexBufferSize;
void (Kobayashi::*create)(CB &) = &Kobayashi::iCreateObject;
void (Kobayashi::*colour)(CB &, int) = &Kobayashi::iUpdateConstantBuffer;
void (Kobayashi::*setapp)(DX12App *) = &Kobayashi::iSetDXApp;
unsigned int (Kobayashi::*cbsize)() = &Kobayashi::iGetConstantBufferSize;
DirectX::XMINT3 (Kobayashi::*count)() = &Kobayashi::iGetObjectCount;
DirectX::XMFLOAT3 (Kobayashi::*osize)() = &Kobayashi::iGetObjectSize;
DirectX::XMFLOAT3 (Kobayashi::*offset)() = &Kobayashi::iGetObjectPositionOffset;
bool (Kobayashi::*updated)() = &Kobayashi::iIsUpdated;
// the viewer holds the simulation through the interface
void drive(ISimulation &s, CB &cb)
{
    s.iCreateObject(cb);
    if (s.iIsUpdated())
        s.iUpdate();
    for (int i = 0; i < (int)cb.size(); ++i)
        s.iUpdateConstantBuffer(cb, i);
    s.iResetSimulationState(cb);
}
"""


def test_reference_signatures_compile(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    src = tmp_path / "signatures.cpp"
    src.write_text(PROGRAM)
    r = subprocess.run([gxx, "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I", CSRC,
                        "-I", INCLUDE, str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_headless_driver_rejects_bad_flags():
    """Argument errors are reported before any GPU call (exit status 2)."""
    binp = os.path.join(ROOT, "crystalgrowth_amd", "bin", "kobayashi_headless")
    if not os.path.exists(binp):
        pytest.fail("headless driver not built (crystalgrowth_amd/build.py)")
    for args in (["--bogus", "1"], ["--delta", "abc"], ["--boundary", "mirror"]):
        r = subprocess.run([binp, *args], capture_output=True, text=True, timeout=60)
        assert r.returncode == 2, (args, r.stderr)
        assert "kobayashi_headless:" in r.stderr
