import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libfwhip.so")


@pytest.fixture(scope="session")
def small_model():
    """8^3-cell box, 2x2x2 subdomains in 2 parts, stretched PML, PEC walls."""
    from adventure_fullwave_amd import mesh
    return mesh.make_config(1, small=(8, (2, 2, 2), (2, 1, 1)))


@pytest.fixture(scope="session")
def cfg1_model():
    from adventure_fullwave_amd import mesh
    return mesh.make_config(1)


@pytest.fixture(scope="session")
def cfg1_sigma_model():
    """cfg 1 with the reference-mode absorber (graded conductor layers)."""
    from adventure_fullwave_amd import mesh
    return mesh.make_config(1, pml_mode="sigma")


@pytest.fixture(scope="session")
def small_sigma_model():
    """The small model with the reference-mode absorber."""
    from adventure_fullwave_amd import mesh
    return mesh.make_config(1, small=(8, (2, 2, 2), (2, 1, 1)), pml_mode="sigma")
