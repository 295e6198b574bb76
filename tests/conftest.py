import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libnsx.so on cuda:0)")


@pytest.fixture(scope="session")
def models(tmp_path_factory):
    """Seeded synthetic sd-shell models in NuShellX file format (generated once per session)."""
    from tests.modelgen import make_models
    return make_models(str(tmp_path_factory.mktemp("models")))


@pytest.fixture(scope="session")
def dops(models):
    """Separable (.dop) operators of the session's models."""
    from tests.modelgen import make_dops
    return make_dops(models)
