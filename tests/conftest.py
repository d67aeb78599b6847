"""Shared fixtures: package loader, oracle adapter, GPU marker."""

import importlib.util
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def load_package():
    """Register `sand-atlas-dem_amd/` as the importable package `sand_atlas_dem_amd`."""
    if "sand_atlas_dem_amd" in sys.modules:
        return sys.modules["sand_atlas_dem_amd"]
    root = ROOT / "sand-atlas-dem_amd"
    spec = importlib.util.spec_from_file_location(
        "sand_atlas_dem_amd", root / "__init__.py", submodule_search_locations=[str(root)]
    )
    module = importlib.util.module_from_spec(spec)
    sys.modules["sand_atlas_dem_amd"] = module
    spec.loader.exec_module(module)
    return module


if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
load_package()
# Every engine packs its templates with the native packer; build it if a fresh
# checkout has not been built yet (plain g++, a few seconds).
if not sys.modules["sand_atlas_dem_amd"].native.PACK_LIBRARY.exists():
    sys.modules["sand_atlas_dem_amd"].native.compile_packer()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP backend)")


@pytest.fixture(scope="session")
def pkg():
    return load_package()


@pytest.fixture(scope="session")
def oracle_module():
    from oracle import oracle

    oracle.build(reference=(Path("/root/reference").exists()))
    return oracle
