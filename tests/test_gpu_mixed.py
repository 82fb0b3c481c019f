"""Mixed-precision form (FP32 reconstruction/projection/wave structure, FP64 state,
pressure recovery, physical fluxes and dt) against the FP64 device path, which the
parity suite pins to the reference.

Tolerances: the headline target is L1(rho) <= 1e-5 against FP64; these runs are short, so
they must stay well inside it (relative L1 <= 1e-7 unless stated).
"""
import numpy as np
import pytest

from overheatingproblem_amd import cases
from overheatingproblem_amd.capi import Solver
from overheatingproblem_amd.config import (DEBUG_LARGE_ARRAY_FORMS, DEBUG_MARCH_FORMS,
                                           DEBUG_MIXED_ALL_FP64, PRECISION_MIXED)

pytestmark = pytest.mark.gpu
GC = 4


def run(cfg, U0, steps=None, dts=None):
    """Advance with the given dt list, or `steps` steps with dt from the reduction."""
    with Solver(cfg) as s:
        s.set_state(U0)
        used = []
        for k in range(len(dts) if dts is not None else steps):
            dt = dts[k] if dts is not None else s.get_delta_t()
            used.append(dt)
            s.time_advance(dt)
        U = s.get_state()
        faults = s.sync()
    return U, np.array(used), faults


def mixed(cfg, **kw):
    c = cfg.copy()
    c.precision = PRECISION_MIXED
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def rel_l1(a, b):
    return float(np.abs(a - b).mean() / np.abs(b).mean())


@pytest.mark.parametrize("forms", [0, DEBUG_MARCH_FORMS])
@pytest.mark.parametrize("scheme,recon,char", [(2, 5, 1), (3, 5, 0), (1, 2, 1), (10, 3, 1),
                                               (2, 1, 0), (3, 3, 0), (11, 5, 1), (22, 5, 1),
                                               (33, 5, 0), (22, 2, 0)])
def test_fp64_recompute_route_is_exact(scheme, recon, char, forms):
    """With every face sent through the mixed form's FP64 recompute list, the result is
    the FP64 path's, bit for bit (tests the list/fix-up machinery), in the row forms (the
    default at this size) and the marching forms (two sweeps or one pass by order)."""
    c, U0 = cases.four_quadrant(70, 51, cases.LAX_LIU_3, x0=0.8, y0=0.8, perturb=1e-3, seed=9,
                                scheme=scheme, recon_order=recon, characteristic=char)
    U64, dts, _ = run(c, U0, steps=3)
    Umx, _, faults = run(mixed(c, debug_flags=DEBUG_MIXED_ALL_FP64 | forms), U0, dts=dts)
    assert faults & 4  # the fallback was used
    assert np.array_equal(U64, Umx)


@pytest.mark.parametrize("scheme", [2, 1, 3, 10, 11, 22, 33])
def test_mixed_close_to_fp64_lax_liu(scheme):
    c, U0 = cases.lax_liu_3(256, scheme=scheme)
    U64, dts, f64 = run(c, U0, steps=60)
    Umx, _, fmx = run(mixed(c), U0, dts=dts)
    assert f64 == 0 and (fmx & ~4) == 0
    assert np.isfinite(Umx).all()
    err = rel_l1(Umx[0, GC:-GC, GC:-GC], U64[0, GC:-GC, GC:-GC])
    assert err < 1e-7, err


def test_mixed_component_weno():
    c, U0 = cases.lax_liu_3(256, characteristic=False)
    U64, dts, _ = run(c, U0, steps=60)
    Umx, _, _ = run(mixed(c), U0, dts=dts)
    assert rel_l1(Umx[0, GC:-GC, GC:-GC], U64[0, GC:-GC, GC:-GC]) < 1e-7


def test_mixed_conserves_on_periodic_box():
    c, U0 = cases.smooth_wave(256, 256, characteristic=True)
    U, _, _ = run(mixed(c), U0, steps=30)
    before = U0[:, GC:-GC, GC:-GC].sum(axis=(1, 2))
    after = U[:, GC:-GC, GC:-GC].sum(axis=(1, 2))
    scale = np.abs(U[:, GC:-GC, GC:-GC]).sum(axis=(1, 2))
    assert np.all(np.abs(after - before) <= 1e-13 * np.maximum(scale, 1.0))


def test_mixed_smooth_wave_accuracy():
    """Smooth advected wave on a coarse grid (truncation error ~1e-6): the mixed form's
    error against the exact solution matches the FP64 form's to 1% (FP32 round-off of
    the reconstructed deviations is ~1e-9 here)."""
    n = 32
    c, U0 = cases.smooth_wave(n, n, characteristic=True, cfl=0.4)
    t_end = 0.1
    with Solver(c) as s:
        s.set_state(U0)
        s.advance_to(t_end, 100000)
        U64 = s.get_state()
    with Solver(mixed(c)) as s:
        s.set_state(U0)
        s.advance_to(t_end, 100000)
        Umx = s.get_state()
    _, Uex = cases.smooth_wave(n, n, t=t_end)
    e64 = np.abs(U64[0, GC:-GC, GC:-GC] - Uex[0, GC:-GC, GC:-GC]).mean()
    emx = np.abs(Umx[0, GC:-GC, GC:-GC] - Uex[0, GC:-GC, GC:-GC]).mean()
    assert e64 > 1e-7
    assert abs(emx / e64 - 1.0) < 1e-2, (e64, emx)


@pytest.mark.parametrize("scheme", [2, 11, 22, 33])
def test_mixed_noh_positivity(scheme):
    """Noh (p = 1e-6) with the safeguard, HLLC and the heat-capacity solvers (the wall
    heating they address): the mixed form stays admissible and within L1(rho) <= 1e-5
    (relative) of the FP64 form over 150 steps."""
    c, U0 = cases.noh(256, positivity=True, scheme=scheme)
    U64, dts, f64 = run(c, U0, steps=150)
    Umx, _, fmx = run(mixed(c), U0, dts=dts)
    assert (f64 & 8) == 0 and (fmx & 8) == 0
    rho, u, v, p = cases.cons_to_prim(Umx[:, GC:-GC, GC:-GC], c.gamma)
    assert np.isfinite(Umx).all() and (rho > 0).all() and (p > 0).all()
    assert rel_l1(Umx[0, GC:-GC, GC:-GC], U64[0, GC:-GC, GC:-GC]) < 1e-5


def test_mixed_double_mach_reflection():
    """DMR with moving boundaries, Roe + positivity fallback: the mixed form stays
    admissible and close to FP64 over a short run at 256 x 64."""
    c, U0 = cases.dmr(256, 64)
    U64, dts, f64 = run(c, U0, steps=40)
    Umx, _, fmx = run(mixed(c), U0, dts=dts)
    rho, u, v, p = cases.cons_to_prim(Umx[:, GC:-GC, GC:-GC], c.gamma)
    assert np.isfinite(Umx).all() and (rho > 0).all() and (p > 0).all()
    assert rel_l1(Umx[0, GC:-GC, GC:-GC], U64[0, GC:-GC, GC:-GC]) < 1e-6


def _row_vs_march_cases():
    noh = cases.noh(256, positivity=True)                       # safeguard + FP64 fix-ups
    quad = cases.four_quadrant(97, 131, cases.LAX_LIU_3, x0=0.8, y0=0.8, perturb=1e-3,
                               seed=4, characteristic=True)     # partial strip, odd rows
    wave = cases.smooth_wave(130, 70, time_order=2, recon_order=3, scheme=3)
    dmr = cases.dmr(192, 48)                                    # DMR sides, positivity
    grav = cases.four_quadrant(64, 190, cases.LAX_LIU_3, source=1, gravity=-0.5,
                               scheme=1, characteristic=False)  # gravity source term
    return {"noh": noh, "quad": quad, "wave": wave, "dmr": dmr, "grav": grav}


@pytest.mark.parametrize("name", ["noh", "quad", "wave", "dmr", "grav"])
def test_buffer_and_large_array_forms_identical(name):
    """The mixed form's marching kernels (fused y sweep in the LDS-staged marching form, x
    sweep with buffer addressing; RPFVM_DEBUG_MARCH_FORMS at these sizes) and the row forms
    (one row per wavefront, global addressing; RPFVM_DEBUG_LARGE_ARRAY_FORMS, the default on
    small grids and for arrays of 4 GB and more) give the same state, bit for bit, with dt
    from the device reduction.  The marching sweep stores only the y fluxes its fix-up reads
    and the safeguard recomputes the rest; the row sweep stores every one and the safeguard
    reads them (Noh: repairs in the first steps)."""
    cfg, U0 = _row_vs_march_cases()[name]
    c = mixed(cfg, debug_flags=DEBUG_MARCH_FORMS)
    Um, dm, fm = run(c, U0, steps=12)
    Ur, dr, fr = run(mixed(cfg, debug_flags=DEBUG_LARGE_ARRAY_FORMS), U0, steps=12)
    assert np.array_equal(dm, dr)
    assert np.array_equal(Um, Ur, equal_nan=True)
    assert fm == fr
