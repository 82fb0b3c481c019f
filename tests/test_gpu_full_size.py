"""Size-independent properties at the headline grid, 4096 x 4096, for both forms: the sizes
the step is measured at run the LDS-staged marching and ring kernels over many strips and
row blocks, which the small parity grids only begin to tile.

* Lax-Liu configuration 3 with outflow on every side is symmetric under
  (x, y, u, v) -> (y, x, v, u); the reference's x and y HLLC branches round differently
  (include/scheme.h:745 vs 895, 777 vs 929), so the FP64 form keeps the symmetry to
  round-off (1e-11 of the field scale) and the mixed form to its FP32 reconstruction
  (1e-5).  A transposed operand, a swapped face direction or a mis-tiled strip breaks it
  at O(1).
* With periodic walls both forms conserve mass, momentum and energy to the round-off of
  the sums (the first and last faces of each line see identical stencils).
* The mixed form stays within its accuracy target of the FP64 form on the same dts:
  relative L1(rho) <= 1e-7 after 10 steps (the headline target is 1e-5 over a run).
* The headline workload itself, Noh 4096^2 with the positivity safeguard: both forms stay
  admissible with no cell left unrepaired, the mixed form within 1e-7 (relative L1(rho))
  of the FP64 form on its dts, and the device-side loop gives the host-driven loop's
  state and clock bit for bit.
* At 11600 x 11600 an array (4 planes) passes 4 GB, so the step runs the forms with 64-bit
  global addressing (one row per wavefront in the fused sweep); the symmetry still holds.
"""
import numpy as np
import pytest

from overheatingproblem_amd import cases
from overheatingproblem_amd.capi import Solver
from overheatingproblem_amd.config import PRECISION_MIXED

pytestmark = pytest.mark.gpu
GC = 4
N = 4096


def advance(cfg, U0, steps, dts=None):
    with Solver(cfg) as s:
        s.set_state(U0)
        used = []
        for k in range(steps):
            dt = dts[k] if dts is not None else s.get_delta_t()
            used.append(dt)
            s.time_advance(dt)
        U = s.get_state()
        faults = s.sync()
    return U, used, faults


@pytest.mark.parametrize("precision,tol", [(0, 1e-11), (PRECISION_MIXED, 1e-5)])
def test_transpose_symmetry_4096(precision, tol):
    c, U0 = cases.lax_liu_3(N)
    c.precision = precision
    U, _, faults = advance(c, U0, 10)
    assert faults & 1 == 0
    I = U[:, GC:-GC, GC:-GC]
    scale = float(np.abs(I).max())
    for a, b in ((0, 0), (1, 2), (2, 1), (3, 3)):
        assert np.abs(I[a] - I[b].T).max() / scale < tol, (a, b)


@pytest.mark.parametrize("precision", [0, PRECISION_MIXED])
def test_periodic_conservation_4096(precision):
    c, U0 = cases.smooth_wave(N, N, characteristic=1)
    c.precision = precision
    U, _, _ = advance(c, U0, 6)
    before = U0[:, GC:-GC, GC:-GC].sum(axis=(1, 2))
    after = U[:, GC:-GC, GC:-GC].sum(axis=(1, 2))
    n = N * N
    assert np.all(np.abs(after - before) <= 1e-12 * n * np.maximum(1.0, np.abs(before) / n))


def test_mixed_close_to_fp64_4096():
    c, U0 = cases.lax_liu_3(N)
    U64, dts, _ = advance(c, U0, 10)
    cm = c.copy()
    cm.precision = PRECISION_MIXED
    Umx, _, _ = advance(cm, U0, 10, dts=dts)
    r64, rmx = U64[0, GC:-GC, GC:-GC], Umx[0, GC:-GC, GC:-GC]
    assert np.isfinite(rmx).all()
    assert float(np.abs(rmx - r64).mean() / np.abs(r64).mean()) <= 1e-7


def test_noh_headline_4096():
    c, U0 = cases.noh(N, positivity=True)
    steps = 8
    U64, dts, f64 = advance(c, U0, steps)
    cm = c.copy()
    cm.precision = PRECISION_MIXED
    Umx, _, fmx = advance(cm, U0, steps, dts=dts)
    for U, f in ((U64, f64), (Umx, fmx)):
        assert f & 8 == 0
        rho, _, _, p = cases.cons_to_prim(U[:, GC:-GC, GC:-GC], c.gamma)
        assert np.isfinite(U).all() and (rho > 0).all() and (p > 0).all()
    r64, rmx = U64[0, GC:-GC, GC:-GC], Umx[0, GC:-GC, GC:-GC]
    assert float(np.abs(rmx - r64).mean() / np.abs(r64).mean()) <= 1e-7
    del Umx
    with Solver(c) as s:  # the same steps through the device-side loop
        s.set_state(U0)
        t, n = s.advance_to(float("inf"), steps)
        Ud = s.get_state()
    assert n == steps and t == sum(dts)
    assert np.array_equal(Ud, U64)


def test_transpose_symmetry_above_4gb():
    n = 11600
    c, U0 = cases.lax_liu_3(n)
    assert 4 * U0[0].size * 8 >= 1 << 32  # the arrays exceed the 32-bit buffer offsets
    U, _, faults = advance(c, U0, 2)
    del U0
    assert faults & 1 == 0
    I = U[:, GC:-GC, GC:-GC]
    scale = float(np.abs(I[0]).max())
    for a, b in ((0, 0), (1, 2), (3, 3)):
        assert np.abs(I[a] - I[b].T).max() / scale < 1e-11, (a, b)
