import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C-ABI on cuda:0)")
    config.addinivalue_line("markers", "slow: full-size cases")


@pytest.fixture(scope="session")
def ctx():
    from gpudbmetalcodegen_amd import engine as E

    if E.device_count() < 1:
        pytest.fail("a GPU test ran without a visible GPU")
    c = E.Context(0)
    cache = ROOT / "gpudbmetalcodegen_amd" / "_cocache"
    if cache.exists():
        c.set_cache_dir(str(cache))
    yield c
    c.close()


def pytest_report_header(config):
    return f"gpudb tests, root={ROOT}, HIP_VISIBLE_DEVICES={os.environ.get('HIP_VISIBLE_DEVICES', '')}"
