import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "cu-reconstruction_amd"
for p in (str(PKG), str(ROOT / "oracle"), str(ROOT / "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (MI355X)")
    config.addinivalue_line("markers", "slow: longer CPU test")


def _ensure_built():
    libs = [PKG / "lib" / "librecon.so", PKG / "lib" / "librecon_scene.so",
            ROOT / "oracle" / "lib" / "liboracle.so"]
    if not all(p.exists() for p in libs):
        from recon.build import build
        build()


_ensure_built()


@pytest.fixture(scope="session")
def cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)
    return torch
