"""Initial conditions (host side, numpy) for the step's workloads.

Each builder returns ``(cfg, U)`` with ``U`` a float64 array in the reference
layout ``[4][R][C]`` (R = nx + 8, C = ny + 8; variables rho, rho*u, rho*v, E).
Primitive-to-conserved conversion follows initEulerconser2D
(include/initialize.h:1389-1405).  Cell centres are x_i = (i + 1/2) dx from the
domain's lower corner, as Mesh_2D (include/mesh.h:151-160).

The reference's own initialisers differ from the measured workloads in ways that
matter (hard-coded end times, a non-deterministic ``rand()`` in the 2-D Noh case,
a Mach-3 inflow in place of the double-Mach-reflection states,
include/initialize.h:900-903,1099-1143), so the workloads are built here from
their published definitions and fed identically to every implementation.
"""
from __future__ import annotations

import math

import numpy as np

from .config import (BC_DMR_SHOCK, BC_DMR_WALL, BC_FIXED_VALUE, BC_INFLOW, BC_OUTFLOW,
                     BC_PERIODICITY, BC_REFLECTION, GHOST, SCHEME_HLLC, make_config)


def prim_to_cons(rho, u, v, p, gamma):
    """initEulerconser2D (include/initialize.h:1389-1405), elementwise."""
    rho = np.asarray(rho, dtype=np.float64)
    u = np.asarray(u, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    p = np.asarray(p, dtype=np.float64)
    E = p / (gamma - 1) + 0.5 * rho * (u * u + v * v)
    return np.stack([rho, u * rho, v * rho, E])


def cons_to_prim(U, gamma):
    """Con_to_Pri_2D (include/initialize.h:1437-1454), elementwise."""
    rho = U[0]
    u = U[1] / U[0]
    v = U[2] / U[0]
    p = (U[3] - 0.5 * rho * (u * u + v * v)) * (gamma - 1)
    return rho, u, v, p


def cell_centres(nx, ny, dx, dy, x0=0.0, y0=0.0, ghosts=True):
    """Cell-centre coordinates on the full [R][C] grid (ghost centres extrapolated)."""
    g = GHOST if ghosts else 0
    i = np.arange(-g, nx + g, dtype=np.float64)
    j = np.arange(-g, ny + g, dtype=np.float64)
    X = x0 + (i + 0.5) * dx
    Y = y0 + (j + 0.5) * dy
    return np.meshgrid(X, Y, indexing="ij")


def interior(A):
    g = GHOST
    return A[..., g:-g, g:-g]


# ----------------------------------------------------------------------------------
def four_quadrant(nx, ny, states, *, x0=0.5, y0=0.5, lx=1.0, ly=1.0, gamma=1.4,
                  perturb=0.0, seed=0, **cfg_kw):
    """2-D four-quadrant Riemann problem (Lax & Liu 1998).  ``states`` maps
    'NE', 'NW', 'SW', 'SE' to primitive (rho, u, v, p).  Cells whose centre
    lies at x >= x0 are east, y >= y0 north.  ``perturb`` adds a seeded relative
    density ripple (symmetry breaking for parity tests)."""
    dx, dy = lx / nx, ly / ny
    X, Y = cell_centres(nx, ny, dx, dy)
    east, north = X >= x0, Y >= y0
    prim = np.zeros((4,) + X.shape)
    for key, mask in (("NE", east & north), ("NW", ~east & north), ("SW", ~east & ~north),
                      ("SE", east & ~north)):
        for k in range(4):
            prim[k][mask] = states[key][k]
    if perturb:
        rng = np.random.default_rng(seed)
        prim[0] *= 1.0 + perturb * rng.uniform(-1.0, 1.0, size=X.shape)
    cfg_kw.setdefault("bc", (BC_OUTFLOW,) * 4)
    cfg = make_config(nx, ny, dx=dx, dy=dy, gamma=gamma, **cfg_kw)
    return cfg, prim_to_cons(*prim, gamma)


# Lax-Liu configuration 3 with the interface at (0.8, 0.8) (the 2048^2 FP64 workload).
LAX_LIU_3 = {"NE": (1.5, 0.0, 0.0, 1.5), "NW": (0.5323, 1.206, 0.0, 0.3),
             "SW": (0.138, 1.206, 1.206, 0.029), "SE": (0.5323, 0.0, 1.206, 0.3)}


def lax_liu_3(n=2048, **cfg_kw):
    cfg_kw.setdefault("scheme", SCHEME_HLLC)
    cfg_kw.setdefault("characteristic", True)
    cfg_kw.setdefault("cfl", 0.5)
    return four_quadrant(n, n, LAX_LIU_3, x0=0.8, y0=0.8, **cfg_kw)


def noh(n=4096, *, gamma=5.0 / 3.0, p0=1.0e-6, rho0=1.0, right_top_bc=BC_OUTFLOW, **cfg_kw):
    """2-D Noh implosion on [0,1]^2: rho = 1, p = p0, unit speed toward the origin.
    Symmetry (reflection) at x = 0 and y = 0; the far walls take ``right_top_bc``
    (outflow by default).  Deterministic: no random perturbation."""
    dx = dy = 1.0 / n
    X, Y = cell_centres(n, n, dx, dy)
    r = np.sqrt(X * X + Y * Y)
    u, v = -X / r, -Y / r
    U = prim_to_cons(np.full_like(X, rho0), u, v, np.full_like(X, p0), gamma)
    cfg_kw.setdefault("scheme", SCHEME_HLLC)
    cfg_kw.setdefault("characteristic", True)
    cfg_kw.setdefault("cfl", 0.4)
    cfg = make_config(n, n, dx=dx, dy=dy, gamma=gamma,
                      bc=(BC_REFLECTION, right_top_bc, BC_REFLECTION, right_top_bc), **cfg_kw)
    return cfg, U


def sedov(n=200, *, gamma=1.4, e0=1.0, p_bg=1.0e-5, r_dep=2.0, **cfg_kw):
    """Sedov-type blast on [-1,1]^2: rho = 1, u = v = 0, p = p_bg, energy e0
    deposited uniformly as pressure in cells with centre radius <= r_dep*dx."""
    dx = dy = 2.0 / n
    X, Y = cell_centres(n, n, dx, dy, x0=-1.0, y0=-1.0)
    r = np.sqrt(X * X + Y * Y)
    dep = r <= r_dep * dx
    inner = dep[GHOST:-GHOST, GHOST:-GHOST]
    n_in = int(inner.sum())
    p = np.full_like(X, p_bg)
    p[dep] = (gamma - 1.0) * e0 / (n_in * dx * dy)
    U = prim_to_cons(np.ones_like(X), np.zeros_like(X), np.zeros_like(X), p, gamma)
    cfg_kw.setdefault("scheme", SCHEME_HLLC)
    cfg_kw.setdefault("characteristic", True)
    cfg_kw.setdefault("cfl", 0.5)
    cfg = make_config(n, n, dx=dx, dy=dy, gamma=gamma, bc=(BC_OUTFLOW,) * 4, **cfg_kw)
    return cfg, U


def double_mach(nx=8192, ny=2048, *, gamma=1.4, **cfg_kw):
    """Mach-10 double Mach reflection on [0,4]x[0,1] (Woodward & Colella 1984) with
    static boundaries (fixed post-shock left and top, reflecting bottom): the
    short-run parity workload.  :func:`dmr` has the moving boundaries."""
    dx, dy = 4.0 / nx, 1.0 / ny
    X, Y = cell_centres(nx, ny, dx, dy)
    post = (8.0, 8.25 * math.cos(math.pi / 6), -8.25 * math.sin(math.pi / 6), 116.5)
    pre = (1.4, 0.0, 0.0, 1.0)
    behind = X < 1.0 / 6.0 + Y / math.sqrt(3.0)
    prim = [np.where(behind, post[k], pre[k]) for k in range(4)]
    U = prim_to_cons(*prim, gamma)
    cons_post = prim_to_cons(*post, gamma)
    cfg_kw.setdefault("scheme", 3)
    cfg_kw.setdefault("characteristic", True)
    cfg_kw.setdefault("cfl", 0.5)
    cfg = make_config(nx, ny, dx=dx, dy=dy, gamma=gamma,
                      bc=(BC_FIXED_VALUE, BC_OUTFLOW, BC_REFLECTION, BC_FIXED_VALUE),
                      fixed_state=(cons_post, cons_post, cons_post, cons_post), **cfg_kw)
    return cfg, U


def dmr(nx=8192, ny=2048, *, gamma=1.4, positivity=True, **cfg_kw):
    """Mach-10 double Mach reflection (Woodward & Colella 1984) on [0,4]x[0,1] with the
    moving boundaries of BC_DoubleMach_2D (include/Boundary_Condition.h:480-660): post-shock
    inflow on the left, outflow on the right, the bottom post-shock up to x = 1/6 and a
    reflecting wall beyond, the top post-shock left of the shock trace
    x_s(t) = 1/6 + (1 + 20 t)/sqrt(3) and pre-shock right of it.  Pre-shock
    (1.4, 0, 0, 1), post-shock (8, 8.25 cos 30, -8.25 sin 30, 116.5); initial shock
    y = sqrt(3)(x - 1/6) as initEulerpri2D_DoubleMachReflection
    (include/initialize.h:890-966).  WENO5-char + Roe, RK3, CFL 0.5, with the positivity
    safeguard (first-order face fallback + HLLE cell repair) as the HLL fallback."""
    dx, dy = 4.0 / nx, 1.0 / ny
    X, Y = cell_centres(nx, ny, dx, dy)
    post = (8.0, 8.25 * math.cos(math.pi / 6), -8.25 * math.sin(math.pi / 6), 116.5)
    pre = (1.4, 0.0, 0.0, 1.0)
    ahead = Y < math.sqrt(3.0) * (X - 1.0 / 6.0)
    prim = [np.where(ahead, pre[k], post[k]) for k in range(4)]
    U = prim_to_cons(*prim, gamma)
    cpost = tuple(float(x) for x in prim_to_cons(*post, gamma))
    cpre = tuple(float(x) for x in prim_to_cons(*pre, gamma))
    cfg_kw.setdefault("scheme", 3)
    cfg_kw.setdefault("characteristic", True)
    cfg_kw.setdefault("cfl", 0.5)
    cfg = make_config(nx, ny, dx=dx, dy=dy, gamma=gamma, positivity=positivity,
                      bc=(BC_INFLOW, BC_OUTFLOW, BC_DMR_WALL, BC_DMR_SHOCK), inflow_state=cpost,
                      fixed_state=(cpost, cpost, cpost, cpost), **cfg_kw)
    cfg.lx = 4.0
    cfg.bc_par[2][0] = 1.0 / 6.0
    for k, v in enumerate((1.0 / 6.0, 1.0, math.sqrt(3.0), 20.0)):
        cfg.bc_par[3][k] = v
    for k in range(4):
        cfg.bc_alt_state[3][k] = cpre[k]
    return cfg, U


def smooth_wave(nx, ny, *, gamma=1.4, amp=0.2, t=0.0, **cfg_kw):
    """Advected density wave rho = 1 + amp sin(2 pi (x + y - 2t)), u = v = 1, p = 1 on a
    periodic unit square (a 2-D version of the precision harness,
    include/initialize.h:36-60)."""
    dx, dy = 1.0 / nx, 1.0 / ny
    X, Y = cell_centres(nx, ny, dx, dy)
    rho = 1.0 + amp * np.sin(2.0 * math.pi * (X + Y - 2.0 * t))
    U = prim_to_cons(rho, np.ones_like(X), np.ones_like(X), np.ones_like(X), gamma)
    cfg = make_config(nx, ny, dx=dx, dy=dy, gamma=gamma, bc=(BC_PERIODICITY,) * 4, **cfg_kw)
    return cfg, U


def precision_wave(nx, *, gamma=1.4, **cfg_kw):
    """Init_Precision (include/initialize.h:36-60, 1609-1629): on [0, 2 pi] x [0, 1] with
    ny = 1, rho = 1 + 0.2 sin((i + 1/2) dx), u = 1, v = 0, p = 1 in the interior cells of
    column GC (every other entry zero, as the reference leaves them), periodic in x and
    outflow in y; exact density 1 + 0.2 sin(x - t), end time 2."""
    R, C = nx + 2 * GHOST, 1 + 2 * GHOST
    lx = 2.0 * math.pi
    dx = lx / nx
    rho = np.zeros((R, C))
    u = np.zeros((R, C))
    p = np.zeros((R, C))
    for i in range(GHOST, R - GHOST):
        rho[i, GHOST] = 1.0 + 0.2 * math.sin((i - GHOST + 0.5) * dx)
        u[i, GHOST] = 1.0
        p[i, GHOST] = 1.0
    U = prim_to_cons(rho, u, np.zeros((R, C)), p, gamma)
    cfg = make_config(nx, 1, dx=dx, dy=1.0, gamma=gamma,
                      bc=(BC_PERIODICITY, BC_PERIODICITY, BC_OUTFLOW, BC_OUTFLOW), **cfg_kw)
    return cfg, U


def random_smooth_state(nx, ny, seed, *, gamma=1.4, amp=0.3, vel=1.0, modes=3):
    """A seeded smooth positive state (sum of random Fourier modes) on the full grid."""
    rng = np.random.default_rng(seed)
    X, Y = cell_centres(nx, ny, 1.0 / nx, 1.0 / ny)
    def field(base, a):
        f = np.full_like(X, base)
        for _ in range(modes):
            kx, ky = rng.integers(1, 4, size=2)
            ph = rng.uniform(0, 2 * math.pi)
            f += a * rng.uniform(-1, 1) * np.sin(2 * math.pi * (kx * X + ky * Y) + ph)
        return f
    rho = np.abs(field(1.0, amp / modes)) + 0.2
    p = np.abs(field(1.0, amp / modes)) + 0.2
    u = field(0.0, vel / modes)
    v = field(0.0, vel / modes)
    return prim_to_cons(rho, u, v, p, gamma)


# Quadrant-state ranges of the seeded batch workload: rho ~ U[0.1, 2], p ~ U[0.02, 2],
# u, v ~ U[-1.5, 1.5].
BATCH_RANGES = {"rho": (0.1, 2.0), "p": (0.02, 2.0), "uv": (-1.5, 1.5)}


def batch_problems(count=64, seed=20240611, sizes=range(256, 1537, 64), ranges=BATCH_RANGES):
    """The uneven batch of seeded four-quadrant problems: n uniform over ``sizes``,
    quadrant states drawn from ``ranges`` (rho, u, v, p per quadrant NE, NW, SW, SE),
    interface at (0.5, 0.5), gamma 1.4, solver uniform over {HLLC, HLL, Roe}.  Returns a
    list of (n, states, scheme).  These ranges reach near-vacuum rarefactions, where the
    reference's HLL (wave speeds not positivity bounded) meets negative states within a
    few dozen steps: :func:`batch_case` turns the positivity safeguard on."""
    rng = np.random.default_rng(seed)
    sizes = list(sizes)
    (r0, r1), (p0, p1), (v0, v1) = ranges["rho"], ranges["p"], ranges["uv"]
    out = []
    for _ in range(count):
        n = int(sizes[rng.integers(0, len(sizes))])
        states = {}
        for key in ("NE", "NW", "SW", "SE"):
            states[key] = (float(rng.uniform(r0, r1)), float(rng.uniform(v0, v1)),
                           float(rng.uniform(v0, v1)), float(rng.uniform(p0, p1)))
        scheme = (2, 1, 3)[int(rng.integers(0, 3))]
        out.append((n, states, scheme))
    return out


def batch_case(count=64, seed=20240611, sizes=range(256, 1537, 64), *, characteristic=True,
               cfl=0.5, positivity=True, ranges=BATCH_RANGES, **cfg_kw):
    """Configs and initial states of :func:`batch_problems` (outflow on every side, the
    positivity safeguard on by default): returns (cfgs, states)."""
    cfgs, states = [], []
    for n, quad, scheme in batch_problems(count, seed, sizes, ranges):
        c, U = four_quadrant(n, n, quad, scheme=scheme, characteristic=characteristic, cfl=cfl,
                             positivity=positivity, **cfg_kw)
        cfgs.append(c)
        states.append(U)
    return cfgs, states
