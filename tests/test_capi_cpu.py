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
    assert ctypes.sizeof(Config) == 14 * 4 + 5 * 8 + 4 * 8 + 16 * 8 + 8 + 16 * 8 + 16 * 8


def test_version(lib):
    assert capi.version().startswith("rpfvm-gfx950")


def test_make_config_roundtrip():
    from overheatingproblem_amd.config import config_from_dict
    c = make_config(10, 12, scheme=3, characteristic=True, bc=(1, 0, 2, 3), positivity=True)
    d = config_from_dict(c.as_dict())
    assert c.as_dict() == d.as_dict()


def test_bench_noh_metrics_on_oracle_states(oracle):
    """bench.py's accuracy report runs on small states (the oracle stands in for the
    device forms here; both entries identical -> zero mixed-vs-FP64 error)."""
    import sys
    sys.path.insert(0, ROOT)
    import bench
    from overheatingproblem_amd import cases
    cfg, U = cases.noh(32, positivity=True)
    for _ in range(3):
        oracle.time_advance(cfg, U, oracle.delta_t(cfg, U))
    out = bench.noh_metrics({"fp64": U, "mixed": U.copy()}, 32, cfg.gamma)
    assert out["mixed_vs_fp64"]["l1_rho"] == 0.0
    assert 0.0 < out["fp64"]["rho_plateau_0.1_0.15"] < 20.0


@pytest.mark.parametrize("scheme,precision,message", [
    (5, 0, "Marquina"), (5, 1, "Marquina"),  # the reference's stub solver
])
def test_create_rejects_unsupported_forms(lib, scheme, precision, message):
    """Configuration checks run before any device call, so they are testable here."""
    cfg = make_config(16, 16, scheme=scheme, precision=precision)
    with pytest.raises(capi.RpfvmError, match=message):
        capi.Solver(cfg)


@pytest.mark.parametrize("key,value,message", [
    ("precision", 0, "recon_order, characteristic, time_order, precision"),
    ("time_order", 2, "recon_order, characteristic, time_order, precision"),
])
def test_batch_rejects_mixed_schemes_of_integration(lib, key, value, message):
    """Every problem of a batch shares the reconstruction, time order and precision; a
    second problem that differs is rejected before any device call."""
    cfgs = [make_config(16, 16, scheme=2, precision=1), make_config(16, 16, scheme=11, precision=1)]
    setattr(cfgs[1], key, value)
    with pytest.raises(capi.RpfvmError, match=message):
        capi.Batch(cfgs)
