"""Shared pytest configuration.

``gpu`` marks tests that need an MI355X (they call the HIP library through its
C-ABI); everything else runs on the CPU.
"""

from __future__ import annotations

import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long CPU oracle runs")
