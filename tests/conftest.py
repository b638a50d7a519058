import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libhjoin.so on the GPU)")


@pytest.fixture(scope="session")
def hjo():
    return importlib.import_module("hash-join-optimizer_amd")


@pytest.fixture(scope="session")
def oracle():
    import hjtest
    return hjtest.Oracle()


@pytest.fixture(scope="session")
def ref():
    import hjtest
    r = hjtest.load_ref()
    if r is None:
        pytest.skip("oracle/_ref/libhjref.so not built (reference tree absent)")
    return r


@pytest.fixture(scope="session")
def gpu_ctx(hjo):
    ctx = hjo.Context(arena_bytes=8 << 30)
    yield ctx
    ctx.close()
