"""Shared test setup: markers, the product package and the oracle.

`-m "not gpu"` tests run in this container (no GPU); `-m gpu` tests need an
MI355X and the built libbmi_j2k.so."""
import importlib.util
import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
PKG_DIR = ROOT / "jpeg2000-decoder_amd"
sys.path.insert(0, str(ROOT / "oracle"))


def load_package():
    """Import jpeg2000-decoder_amd/ (the directory name is not an identifier)."""
    name = "jpeg2000_decoder_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(
        name, PKG_DIR / "__init__.py", submodule_search_locations=[str(PKG_DIR)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")


@pytest.fixture(scope="session")
def pkg():
    return load_package()


@pytest.fixture(scope="session")
def workloads(pkg):
    import importlib
    return importlib.import_module("jpeg2000_decoder_amd.workloads")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle
