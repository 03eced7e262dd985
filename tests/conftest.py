"""Shared test setup: the ``gpu`` marker, the repo root on sys.path, and a
fixture that builds the CPU oracle (the checker) once per session."""

from __future__ import annotations

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle as O
    if not O.LIB_PATH.exists():
        O.build()
    return O


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP library; a GPU test must fail loudly (not skip) without it."""
    from roundabout_amd.rdb import _ffi
    return _ffi.load()
