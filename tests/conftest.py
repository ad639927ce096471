import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


def pytest_collection_modifyitems(config, items):
    # GPU tests need a device; on a CPU-only host they are skipped unless the
    # caller explicitly selected them with -m gpu (then they fail loudly).
    try:
        import torch

        has_gpu = torch.cuda.is_available()
    except Exception:  # pragma: no cover - torch missing
        has_gpu = False
    markexpr = config.getoption("-m") or ""
    if has_gpu or "gpu" in markexpr.replace("not gpu", ""):
        return
    skip = pytest.mark.skip(reason="no HIP device on this host")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)
