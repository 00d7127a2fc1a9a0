"""Shared test configuration: the ``gpu`` marker and repository imports."""

from __future__ import annotations

import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):  # type: ignore[no-untyped-def]
    config.addinivalue_line("markers", "gpu: needs a visible MI355X (HIP device); runs the gfx950 kernels")
