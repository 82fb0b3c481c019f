"""Host logic of the convergence harness (overheatingproblem_amd.precision): the norms,
the orders and the table against closed forms and the reference program's own printed
output (tests/golden/driver.npz, Precision_Test runs).  No device work."""
import math
import re

import numpy as np
import pytest

from golden_util import load
from overheatingproblem_amd import cases, precision

GC = 4


def test_orders_of_power_law_errors():
    """e_n = C n^-p on doubling grids gives order p exactly (to round-off)."""
    rows = [dict(n=n, l1=3.0 * n ** -5.0, l2=0.5 * n ** -4.0, linf=n ** -3.5) for n in precision.NX]
    precision.orders(rows)
    for r in rows[1:]:
        assert abs(r["o1"] - 5.0) < 1e-12 and abs(r["o2"] - 4.0) < 1e-12
        assert abs(r["oinf"] - 3.5) < 1e-12


def test_orders_zero_error_gives_zero():
    rows = precision.orders([dict(n=10, l1=0.0, l2=1.0, linf=1.0), dict(n=20, l1=1.0, l2=0.0, linf=0.5)])
    assert rows[1]["o1"] == 0.0 and rows[1]["o2"] == 0.0 and rows[1]["oinf"] == 1.0


def test_density_errors_of_exact_state_and_of_a_shift():
    """Zero for the exact density at t; a uniform offset d gives L1 = 2 pi d,
    L2 = sqrt(2 pi) d, Linf = d (the domain length is 2 pi)."""
    nx, t, d = 64, 0.7, 1e-3
    cfg, U = cases.precision_wave(nx)
    dx = cfg.dx
    for i in range(GC, GC + nx):
        U[0, i, GC] = 1.0 + 0.2 * math.sin((i - GC + 0.5) * dx - t)
    assert precision.density_errors(U, dx, t) == (0.0, 0.0, 0.0)
    U[0, GC:GC + nx, GC] -= d
    l1, l2, linf = precision.density_errors(U, dx, t)
    assert l1 == pytest.approx(2 * math.pi * d, rel=1e-9)
    assert l2 == pytest.approx(math.sqrt(2 * math.pi) * d, rel=1e-9)
    assert linf == pytest.approx(d, rel=1e-9)


def test_step_times_reach_the_end_time():
    for nx in precision.NX:
        dx = 2 * math.pi / nx
        dts, t = precision.step_times(dx)
        assert all(dt == precision.fixed_dt(dx) for dt in dts[:-1])
        assert 0 < dts[-1] <= precision.fixed_dt(dx) and t >= precision.T_END - 1e-12


def test_precision_wave_matches_the_reference_initialiser():
    """Only column GC's interior is set (include/initialize.h:52-57); E = p/(g-1) + rho u^2/2."""
    cfg, U = cases.precision_wave(10, gamma=1.4)
    assert cfg.ny == 1 and tuple(cfg.bc) == (2, 2, 0, 0)
    assert np.count_nonzero(U[0]) == 10 and np.all(U[0, GC:-GC, GC] > 0)
    assert U[3, GC, GC] == 1.0 / (1.4 - 1) + 0.5 * U[0, GC, GC] * (1.0 * 1.0 + 0.0)


@pytest.mark.parametrize("name", ["precision_hllc", "precision_hllhc", "precision_roehc_rk2"])
def test_table_and_orders_match_the_reference_printout(name):
    """From the errors the reference printed (%.10e), orders() and format_table()
    reproduce its order lines and its error table character for character."""
    out = str(load("driver")[name + "__stdout"])
    num = r"([0-9.e+-]+)"
    l1 = [float(x) for x in re.findall(r"L1 Error = " + num, out)]
    l2 = [float(x) for x in re.findall(r"L2 Error = " + num, out)]
    li = [float(x) for x in re.findall("L∞ Error = " + num, out)]
    rows = precision.orders([dict(n=n, dx=2 * math.pi / n, l1=a, l2=b, linf=c)
                             for n, a, b, c in zip(precision.NX, l1, l2, li)])
    for k, r in enumerate(rows[1:], 1):
        line = (f"Order calculation for grid {k - 1} (N={rows[k - 1]['n']}) to {k} (N={r['n']}): "
                f"h_ratio = 2.0000, L1_order = {r['o1']:.4f}")
        assert line in out
    table = precision.format_table(rows)
    assert table in out
