import os
import sys
import pathlib

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libpcs.so")


@pytest.fixture(scope="session")
def gpu_lib():
    """libpcs.so on a visible GPU. GPU tests fail (not skip) when the library
    or the device is missing: there is no other compute path."""
    from part_conc_spt_abc_amd import _abi

    lib = _abi.load()
    n = lib.pcs_device_count()
    assert n > 0, "no GPU visible to libpcs.so"
    return lib
