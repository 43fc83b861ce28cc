"""The oracle pins catch plausible mistakes -- CPU only.

Each mutant is the oracle source compiled with -DKGO_MUTATION=k (one planted
mistake, see oracle/kobayashi_oracle_impl.inc); the pin suite is then run
against it in a subprocess and must fail.  Nothing here touches a GPU.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")

MUTANTS = {
    1: "t2 sign flipped",
    2: "t1 reads gy instead of gx",
    3: "first differences over 2dx",
    4: "t3 (grad eps^2 . grad phi) dropped",
    5: "Laplacian diagonals dropped",
    6: "zero-flux clamp in pass 1 only",
    7: "T update with K (phi - phi') (sign)",
    8: "theta branch gx>0, gy<0 missing 2 pi",
    9: "eps' sign flipped",
    10: "m(T) uses T - Teq",
    11: "noise outside the p(1-p) factor",
    13: "periodic west index off by one",
    14: "theta0 offset with wrong sign",
}


def _build(k, out):
    cmd = ["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-std=c99", "-shared",
           f"-DKGO_MUTATION={k}", "-o", out, os.path.join(ORACLE, "kobayashi_oracle.c"), "-lm"]
    subprocess.run(cmd, check=True)


@pytest.mark.parametrize("k", sorted(MUTANTS))
def test_pins_catch_mutant(k, tmp_path):
    lib = str(tmp_path / f"libkg_oracle_m{k}.so")
    _build(k, lib)
    env = dict(os.environ, KGO_ORACLE_LIB=lib)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", "test_oracle_pins.py")],
                       env=env, cwd=ROOT, capture_output=True, text=True, timeout=300)
    assert r.returncode != 0, f"mutant {k} ({MUTANTS[k]}) passed every pin:\n{r.stdout[-2000:]}"
    assert "failed" in r.stdout
