import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def lib():
    from spral_amd import _lib
    if not _lib.LIB_PATH.exists():
        from spral_amd import build
        build.build(verbose=False)
    return _lib.lib()


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle_py
    if not (oracle_py.HERE / "liboracle.so").exists():
        from spral_amd import build
        build.build_oracle()
    return oracle_py.lib()
