"""Shared pytest setup: the `gpu` marker and the built libraries."""
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built libumi_hip.so")


@pytest.fixture(scope="session")
def gpu():
    """Initialises the device once; fails (does not skip) when there is none."""
    import umi_amd

    n = umi_amd.device_count()
    assert n >= 1, "no HIP device visible: GPU tests must run on the MI355X box"
    print(f"[gpu] {umi_amd.device_name()}", flush=True)
    return umi_amd
