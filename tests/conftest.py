import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("gpu-path-tracer_amd", "oracle"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on the GPU)")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def cornell():
    import scenes
    return scenes.cornell_box()


@pytest.fixture(scope="session")
def cornell_oracle(cornell, oracle_lib):
    return oracle_lib.OracleScene(cornell)


@pytest.fixture(scope="session")
def gpu_renderer():
    import ptcore
    r = ptcore.Renderer()
    yield r
    r.close()
