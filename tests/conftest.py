import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def sg():
    import dune_sculpt_amd as sg

    sg.lib()
    return sg


@pytest.fixture(scope="session")
def po():
    import pyoracle as po

    po.lib()
    return po
