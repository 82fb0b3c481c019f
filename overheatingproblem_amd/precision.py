"""Grid-convergence harness of the step on the GPU (the reference's Precision_test,
src/RP_FVM.c:974-1072, with the norms and table of include/Error.h:129-329).

The advected sine wave of Init_Precision (``cases.precision_wave``) runs on nx = 10 ..
640 cells to T = 2 with the fixed dt = 0.5 dx^(5/3) of Precision_Get_Delta_T_2D
(include/function.h:239-241); the L1, L2 and Linf errors of the density along column
GC against 1 + 0.2 sin(x - t) and the orders between successive grids are returned and
printed as the reference's table.  Any configuration of the library runs through it --
solver, reconstruction, characteristic mode, RK order and the FP64 or mixed form -- so
the same harness is the convergence regression of every form.

    python -m overheatingproblem_amd.precision [--scheme 2] [--time-order 3]
        [--recon 5] [--characteristic] [--mixed] [--nx 10,20,...]

The steps run on the device through the C ABI; the host computes only the norms (a
few hundred cells), in the reference's summation order, with libm's sin.
"""
from __future__ import annotations

import argparse
import math

from . import cases
from .capi import Solver
from .config import PRECISION_FP64, PRECISION_MIXED

NX = (10, 20, 40, 80, 160, 320, 640)
T_END = 2.0
GC = 4


def fixed_dt(dx: float) -> float:
    """Precision_Get_Delta_T_2D (include/function.h:239-241)."""
    return 0.5 * math.pow(dx, 5.0 / 3)


def step_times(dx: float, t_end: float = T_END):
    """The dts of the harness loop (src/RP_FVM.c:1023-1049): fixed dt, the last one
    clipped to reach t_end; returns (dts, t) with t the loop's final clock."""
    dts, t = [], 0.0
    while t < t_end:
        dt = fixed_dt(dx)
        if t + dt >= t_end:
            dt = t_end - t
        dts.append(dt)
        t = t + dt
    return dts, t


def density_errors(U, dx: float, t: float):
    """TEST_Scheme_Error_L1 / L2 / Linf (include/Error.h:129-242): sums in cell order."""
    R = U.shape[1]
    l1 = l2 = linf = 0.0
    for i in range(GC, R - GC):
        err = 1.0 + 0.2 * math.sin((i - GC + 0.5) * dx - t) - float(U[0, i, GC])
        l1 += abs(err) * dx
        l2 += err * err * dx
        if abs(err) > linf:
            linf = abs(err)
    return l1, math.sqrt(l2), linf


def orders(rows):
    """calculate_orders (include/Error.h:259-294): ln(e_coarse / e_fine) / ln(N_fine /
    N_coarse), 0 where an error is not positive; fills o1, o2, oinf from the second row."""
    for a, b in zip(rows, rows[1:]):
        h = float(b["n"]) / float(a["n"])
        for e, o in (("l1", "o1"), ("l2", "o2"), ("linf", "oinf")):
            b[o] = math.log(a[e] / b[e]) / math.log(h) if a[e] > 0 and b[e] > 0 else 0.0
    return rows


def run_grid(nx: int, **cfg_kw):
    """One grid on the device: returns (U after the loop, final clock, steps, faults)."""
    cfg, U0 = cases.precision_wave(nx, **cfg_kw)
    dts, t = step_times(cfg.dx)
    with Solver(cfg) as s:
        s.set_state(U0)
        for dt in dts:
            s.time_advance(dt)
        faults = s.sync()
        U = s.get_state()
    return U, t, len(dts), faults


def convergence(nx_values=NX, **cfg_kw):
    """The harness over nx_values; cfg_kw go to the configuration (scheme, time_order,
    recon_order, characteristic, precision, gamma).  Returns one dict per grid."""
    rows = []
    for nx in nx_values:
        U, t, steps, faults = run_grid(nx, **cfg_kw)
        dx = 2.0 * math.pi / nx
        l1, l2, linf = density_errors(U, dx, t)
        rows.append(dict(n=nx, dx=dx, l1=l1, l2=l2, linf=linf, steps=steps, faults=faults))
    return orders(rows)


def format_table(rows) -> str:
    """print_error_table (include/Error.h:303-329)."""
    rule = "=" * 79
    out = ["", "", rule, "Table 3.3: Errors and numerical orders of accuracy of the density", rule,
           "Grid points  Δx        L1 error    Order     L2 error    Order     L∞ error    Order",
           "-" * 78]
    r = rows[0]
    out.append(f"{r['n']:<11d}  {r['dx']:<9.6f}  {r['l1']:<10.2e}  {'---':<8s}  {r['l2']:<10.2e}  "
               f"{'---':<8s}  {r['linf']:<10.2e}  {'---':<8s}")
    for r in rows[1:]:
        out.append(f"{r['n']:<11d}  {r['dx']:<9.6f}  {r['l1']:<10.2e}  {r['o1']:<8.2f}  "
                   f"{r['l2']:<10.2e}  {r['o2']:<8.2f}  {r['linf']:<10.2e}  {r['oinf']:<8.2f}")
    out.append(rule)
    return "\n".join(out) + "\n"


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--scheme", type=int, default=2)
    ap.add_argument("--time-order", type=int, default=3)
    ap.add_argument("--recon", type=int, default=5)
    ap.add_argument("--characteristic", action="store_true")
    ap.add_argument("--mixed", action="store_true")
    ap.add_argument("--gamma", type=float, default=1.4)
    ap.add_argument("--nx", default=",".join(map(str, NX)))
    a = ap.parse_args(argv)
    rows = convergence([int(x) for x in a.nx.split(",")], scheme=a.scheme,
                       time_order=a.time_order, recon_order=a.recon,
                       characteristic=a.characteristic, gamma=a.gamma,
                       precision=PRECISION_MIXED if a.mixed else PRECISION_FP64)
    for r in rows:
        print(f"Nx = {r['n']}: L1 {r['l1']:.10e}  L2 {r['l2']:.10e}  Linf {r['linf']:.10e}  "
              f"({r['steps']} steps, fault bits {r['faults']})")
    print(format_table(rows), end="")


if __name__ == "__main__":
    main()
