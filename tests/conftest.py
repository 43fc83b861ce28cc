"""Shared pytest configuration.

Markers:
  gpu -- needs an MI355X (runs through the HIP C-ABI library).
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: test needs a GPU (MI355X) and the HIP library")


@pytest.fixture(scope="session")
def O():
    """The CPU oracle binding (oracle/oracle.py)."""
    from oracle import oracle
    oracle.lib()
    return oracle
