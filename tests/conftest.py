"""Shared fixtures.  GPU tests are marked `gpu` and skip cleanly without a device."""
from __future__ import annotations

import ctypes
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")
    config.addinivalue_line("markers", "slow: full-size workload")


def _gpu_count() -> int:
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            hip = ctypes.CDLL(name)
        except OSError:
            continue
        n = ctypes.c_int(0)
        if hip.hipGetDeviceCount(ctypes.byref(n)) == 0:
            return n.value
        return 0
    return 0


_GPUS = None


def gpu_available() -> bool:
    global _GPUS
    if _GPUS is None:
        _GPUS = _gpu_count()
    return _GPUS > 0


@pytest.fixture(autouse=True)
def _skip_without_gpu(request):
    if request.node.get_closest_marker("gpu") and not gpu_available():
        pytest.skip("no HIP device")


@pytest.fixture(autouse=True)
def _announce(request):
    # name the case before any device work, so a crash log still identifies it
    if request.node.get_closest_marker("gpu"):
        print(f"\n[gpu case] {request.node.nodeid}", flush=True)
    yield
