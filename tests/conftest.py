import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "msbg-flip_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def has_gpu():
    """A ROCm device node is present.  Deliberately not torch.cuda: torch
    carries its own HIP runtime, and initialising it in the same process as
    libmsbgx's runtime leaves the second one without a device."""
    return os.path.exists("/dev/kfd")
