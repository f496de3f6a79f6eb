import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(REPO / "webgpu-3dgs_amd"))
sys.path.insert(0, str(REPO / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built libgs3d.so")


@pytest.fixture(scope="session")
def renderer():
    from gs3d.renderer import GaussianSplatRenderer

    r = GaussianSplatRenderer(0)
    yield r
    r.close()
