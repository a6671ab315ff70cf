import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT / "smearfem.jl_amd", ROOT / "oracle", ROOT):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def ctx():
    import smearfem

    c = smearfem.Context(0)
    info = c.device_info()
    print(f"\n[gpu] {info}", flush=True)
    yield c
    c.close()
