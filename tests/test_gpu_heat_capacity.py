"""Riemann solvers alone and the heat-capacity schemes (11 HLLHC, 22 HLLCHC, 33 RoeHC)
on the device, FP64 form, through the C ABI.

Pinned bit-for-bit to the reference's own flux functions and steps (tests/golden/
flux.npz, flux_hc.npz, written by the reference's code) and to the CPU oracle on a
solver x reconstruction x mode matrix.  Equality is exact (+0 == -0, NaN matches NaN).
"""
import numpy as np
import pytest

from golden_util import cfg as golden_cfg, keys_with_suffix, load, same
from overheatingproblem_amd import cases
from overheatingproblem_amd.capi import Solver
from overheatingproblem_amd.config import DEBUG_MARCH_FORMS

pytestmark = pytest.mark.gpu
GC = 4

FLUX = load("flux")
HC = load("flux_hc")


def xfaces(c):
    return (slice(None), slice(GC - 1, c.R - GC), slice(GC, c.C - GC))


def yfaces(c):
    return (slice(None), slice(GC, c.R - GC), slice(GC - 1, c.C - GC))


def first_diff(a, b):
    bad = ~((a == b) | (np.isnan(a) & np.isnan(b)))
    idx = np.argwhere(bad)
    if len(idx) == 0:
        return "identical"
    k = tuple(idx[0])
    return f"{len(idx)} differ; first at {k}: {a[k]!r} vs {b[k]!r}"


def skipped_faces(c, Ls, Rs):
    """Faces the reference skips (leaving its flux array as it was; the harness that
    wrote the fixtures zero-fills it): RoeHC with a NaN Roe density sqrt(rho_L rho_R),
    include/scheme.h:2936-2939.  The device writes NaN there and raises fault bit 1."""
    if c.scheme != 33:
        return np.zeros(Ls.shape[1:], dtype=bool)
    with np.errstate(invalid="ignore"):
        return np.isnan(np.sqrt(Ls[0]) * np.sqrt(Rs[0]))


def riemann_case(d, key):
    c = golden_cfg(d, key + "_cfg")
    Ls, Rs = np.array(d[key + "_L"]), np.array(d[key + "_R"])
    with Solver(c) as s:
        F1 = s.riemann(1, Ls, Rs)
        F2 = s.riemann(2, Ls, Rs)
        faults = s.sync()
    skip = skipped_faces(c, Ls, Rs)
    for F, ref, faces in ((F1, d[key + "_F1"], xfaces(c)), (F2, d[key + "_F2"], yfaces(c))):
        want = np.where(skip[None], np.nan, ref)[faces]
        assert same(F[faces], want), first_diff(F[faces], want)
    assert bool(faults & 1) == bool(skip[xfaces(c)[1:]].any() or skip[yfaces(c)[1:]].any())


@pytest.mark.parametrize("key", keys_with_suffix(FLUX, "_cfg"))
def test_riemann_golden(key):
    """HLL, HLLC, Roe, LLF on the reference's flux fixtures."""
    riemann_case(FLUX, key)


@pytest.mark.parametrize("key", [k for k in keys_with_suffix(HC, "_cfg") if not k.startswith("step_")])
def test_riemann_heat_capacity_golden(key):
    """HLLHC, HLLCHC, RoeHC on physical, supersonic, resting, near-vacuum,
    negative-pressure and negative-density faces."""
    riemann_case(HC, key)


@pytest.mark.parametrize("key", [k for k in keys_with_suffix(HC, "_cfg") if k.startswith("step_")])
def test_golden_steps_heat_capacity(key):
    c = golden_cfg(HC, key + "_cfg")
    ref_dts = HC[key + "_dts"]
    with Solver(c) as s:
        s.set_state(HC[key + "_U0"])
        dts = []
        for _ in range(len(ref_dts)):
            dts.append(s.get_delta_t())
            s.time_advance(dts[-1])
        U = s.get_state()
        assert s.sync() == 0
    assert same(np.array(dts), ref_dts), (dts, ref_dts)
    assert same(U, HC[key + "_U"]), first_diff(U, HC[key + "_U"])


MATRIX = [(s, r, ch, o) for s in (11, 22, 33) for r in (5, 3, 2, 1) for ch in (1, 0)
          for o in ((3, 2, 1) if r == 5 else (3,))]


@pytest.mark.parametrize("forms", ["row", "march"])
@pytest.mark.parametrize("scheme,recon,char,order", MATRIX)
def test_heat_capacity_oracle_matrix(oracle, scheme, recon, char, order, forms):
    """Every heat-capacity solver x reconstruction x mode (and RK order for WENO5) on a
    non-square grid whose sizes are not multiples of the 64-lane tile, 3 steps, in the row
    forms (the default at this size) and the marching forms."""
    c, U0 = cases.four_quadrant(67, 45, cases.LAX_LIU_3, x0=0.8, y0=0.8, perturb=1e-3,
                                seed=scheme * 100 + recon * 10 + char, scheme=scheme,
                                recon_order=recon, characteristic=char, time_order=order)
    if forms == "march":
        c.debug_flags = DEBUG_MARCH_FORMS
    U = U0.copy()
    ref_dts = oracle.run_steps(c, U, 3)
    with Solver(c) as s:
        s.set_state(U0)
        dts = []
        for _ in range(3):
            dts.append(s.get_delta_t())
            s.time_advance(dts[-1])
        G = s.get_state()
    assert same(np.array(dts), ref_dts)
    assert same(G, U), first_diff(G, U)


@pytest.mark.parametrize("scheme", [11, 22, 33])
def test_heat_capacity_sedov_256_prefix(oracle, scheme):
    """A 256^2 Sedov blast, 4 steps of the device loop (advance_to), against the
    oracle's loop."""
    c, U0 = cases.sedov(256, scheme=scheme)
    U = U0.copy()
    t_ref, dts = oracle.advance_to(c, U, 1.0, 4)
    with Solver(c) as s:
        s.set_state(U0)
        t, n = s.advance_to(1.0, 4)
        G = s.get_state()
    assert n == len(dts) == 4 and t == t_ref
    assert same(G, U), first_diff(G, U)
