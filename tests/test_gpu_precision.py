"""The convergence harness (Precision_test) on the device, through the C ABI.

The FP64 form's states after the harness loop equal the oracle's bit for bit, so its
errors are the reference's; WENO5 with SSP-RK3 and dt ~ dx^(5/3) converges at order 5
(the design order; the reference's printed table shows 5.00-5.01); the mixed form
follows the FP64 errors until its FP32 face arithmetic sets a floor.
"""
import numpy as np
import pytest

from golden_util import same
from overheatingproblem_amd import cases, precision
from overheatingproblem_amd.config import PRECISION_MIXED

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scheme,order", [(2, 3), (11, 3), (33, 2), (3, 1)])
def test_harness_states_match_oracle(oracle, scheme, order):
    for nx in (10, 20, 40):
        U, t, steps, faults = precision.run_grid(nx, scheme=scheme, time_order=order)
        cfg, V = cases.precision_wave(nx, scheme=scheme, time_order=order)
        dts, t_ref = precision.step_times(cfg.dx)
        for dt in dts:
            oracle.time_advance(cfg, V, dt)
        assert faults == 0 and t == t_ref and steps == len(dts)
        assert same(U, V), (nx, float(np.nanmax(np.abs(U - V))))


def test_fp64_fifth_order():
    rows = precision.convergence((40, 80, 160, 320, 640))
    for r in rows[1:]:
        assert 4.8 < r["o1"] < 5.3 and 4.8 < r["o2"] < 5.3, precision.format_table(rows)
    assert rows[-1]["l1"] < 3e-11


def test_mixed_form_convergence():
    """Mixed form: the FP64 errors where truncation dominates (to 1e-3 relative up to
    nx = 160; measured 3e-4), fifth order there, and a floor set by its FP32 face
    arithmetic on the finest grid (measured L1 4.2e-11 at nx = 640, FP64 1.8e-11)."""
    fp = precision.convergence(precision.NX)
    mx = precision.convergence(precision.NX, precision=PRECISION_MIXED)
    print(precision.format_table(fp), precision.format_table(mx))
    for a, b in zip(fp, mx):
        print(a["n"], a["l1"], b["l1"], b["l1"] / a["l1"], b["faults"])
    for a, b in zip(fp[:5], mx[:5]):
        assert abs(b["l1"] / a["l1"] - 1.0) < 1e-3 and b["faults"] == 0
    for r in mx[2:5]:
        assert 4.8 < r["o1"] < 5.3
    assert mx[-1]["l1"] < 1e-10
