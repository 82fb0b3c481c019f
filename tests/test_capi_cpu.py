"""CPU-side checks of the C ABI: the library loads and exports every entry point
include/rpfvm.h declares (no device work is done here)."""
import ctypes
import os
import re

import pytest

from overheatingproblem_amd import capi
from overheatingproblem_amd.config import Config, make_config

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"\b(rpfvm_\w+)\s*\(", text))
    return sorted(names)


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(capi.LIB_PATH):
        from overheatingproblem_amd import build
        build.build(verbose=False)
    return capi.lib()


def test_exports_every_declared_symbol(lib):
    names = declared_entry_points()
    assert len(names) >= 16
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(capi.EXPORTS) <= set(names)


def test_config_struct_layout_matches_header(lib):
    # rpfvm_config_default writes through the ctypes struct: layout must agree.
    c = Config()
    lib.rpfvm_config_default(ctypes.byref(c))
    assert (c.nx, c.ny, c.time_order, c.scheme, c.recon_order) == (100, 100, 3, 2, 5)
    assert c.characteristic == 0 and c.positivity == 0 and list(c.bc) == [0, 0, 0, 0]
    assert c.gamma == 1.4 and c.cfl == 0.5 and c.dx == 0.01
    assert c.inflow_state[1] == 1.4 * 3.0 and c.fixed_state[3][3] == 1.0 / 0.4 + 0.5 * 1.4 * 3.0 * 3.0
    assert ctypes.sizeof(Config) == 14 * 4 + 5 * 8 + 4 * 8 + 16 * 8


def test_version(lib):
    assert capi.version().startswith("rpfvm-gfx950")


def test_make_config_roundtrip():
    from overheatingproblem_amd.config import config_from_dict
    c = make_config(10, 12, scheme=3, characteristic=True, bc=(1, 0, 2, 3), positivity=True)
    d = config_from_dict(c.as_dict())
    assert c.as_dict() == d.as_dict()
