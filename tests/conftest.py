"""Shared test set-up: repository on sys.path, the package loader, the
`gpu` marker and the CPU oracle fixtures."""

from __future__ import annotations

import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parents[1]
if str(REPO) not in sys.path:
    sys.path.insert(0, str(REPO))

import __graft_entry__ as entry  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def pkg():
    return entry.load_package()


@pytest.fixture(scope="session")
def scene(pkg):
    return entry.load_submodule("scene")


@pytest.fixture(scope="session")
def ref32():
    from oracle.ref_cpu import RefCpu

    return RefCpu("f32")


@pytest.fixture(scope="session")
def ref64():
    from oracle.ref_cpu import RefCpu

    return RefCpu("f64")
