import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (str(ROOT), str(ROOT / "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs the HIP path on an MI355X (skipped without a HIP device)")


_GPU = None


def gpu_available() -> bool:
    global _GPU
    if _GPU is None:
        from goqu_amd import _native  # raises if libgoqu_hip.so is missing: no silent fallback

        _GPU = _native.device_count() > 0
    return _GPU


def pytest_collection_modifyitems(config, items):
    gpu_items = [it for it in items if "gpu" in it.keywords]
    if not gpu_items:
        return
    if not gpu_available():
        skip = pytest.mark.skip(reason="no HIP device visible")
        for it in gpu_items:
            it.add_marker(skip)
