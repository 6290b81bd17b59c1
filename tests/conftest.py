"""Shared test configuration.

Markers:
  gpu   needs an MI355X (run with ``-m gpu``); every GPU test calls the HIP
        library through its C ABI.
  slow  long CPU oracle checks; skipped unless PORERAD_FULL=1.
"""

from __future__ import annotations

import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PACKAGE_DIR = os.path.join(REPO, "pore-radiation-ijhmt2026_amd")
for path in (REPO, PACKAGE_DIR, os.path.join(REPO, "tests")):
    if path not in sys.path:
        sys.path.insert(0, path)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long CPU oracle check, run with PORERAD_FULL=1")


def pytest_collection_modifyitems(config, items):
    if os.environ.get("PORERAD_FULL") == "1":
        return
    skip = pytest.mark.skip(reason="long oracle check; set PORERAD_FULL=1")
    for item in items:
        if "slow" in item.keywords:
            item.add_marker(skip)
