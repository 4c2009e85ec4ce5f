import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "ferrite.jl_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C-ABI on the device)")
    config.addinivalue_line("markers", "slow: larger CPU cases")


def gpu_available():
    """Whether a device is visible, without initialising torch's HIP runtime (see test_abi)."""
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False
