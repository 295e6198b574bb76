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


@pytest.fixture(scope="session")
def pf44():
    """pf shell, 4 protons + 4 neutrons, J = 0 (N = 41,355; block sides up to 99)."""
    import tempfile
    from nushellx_amd.synth.interaction import random_interaction
    from nushellx_amd.synth.mspace import PF, PNSpace
    from nushellx_amd.synth.writer import PNModel, write_model
    print("generating pf 4p4n J=0 operator files", flush=True)
    inter = random_interaction(PF, seed=2024, spe_range=(-10.0, -4.0))
    return write_model(PNModel(PNSpace(PF), 4, 4, inter), tempfile.mkdtemp(prefix="nsx_pf44_"), "pf44", [0])
