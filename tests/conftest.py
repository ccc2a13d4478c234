"""Test setup: registers the `yaqs-julia_amd` directory as the importable
package `yaqs_julia_amd`, puts the repo root on sys.path (for `oracle`), and
declares the `gpu` marker."""
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "yaqs-julia_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")

if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
if "yaqs_julia_amd" not in sys.modules:
    spec = importlib.util.spec_from_file_location("yaqs_julia_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["yaqs_julia_amd"] = mod
    spec.loader.exec_module(mod)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI library)")
    config.addinivalue_line("markers", "slow: long CPU-oracle runs")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle as O
    O.build()
    return O
