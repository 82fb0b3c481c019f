"""Write the golden fixtures under tests/golden/ from the reference's own code.

Runs oracle/_ref/ref_harness (the reference solver's headers compiled with the
driver in oracle/ref_harness.c; build it with ``make -C oracle ref``) on seeded
inputs and stores inputs and outputs as .npz data.  Run in the build container,
where the reference sources exist:

    python tests/golden/make_golden.py

The fixtures are data only (inputs, outputs, the configuration as JSON).
"""
from __future__ import annotations

import json
import math
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.pyoracle import harness_job  # noqa: E402
from overheatingproblem_amd import cases  # noqa: E402
from overheatingproblem_amd.config import make_config  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
OUT = os.path.dirname(os.path.abspath(__file__))


job_text = harness_job


def run(job, inputs, n_out):
    with tempfile.TemporaryDirectory() as td:
        jp, ip, op = (os.path.join(td, n) for n in ("job.txt", "in.bin", "out.bin"))
        with open(jp, "w") as f:
            f.write(job)
        np.concatenate([np.ravel(a) for a in inputs]).astype("<f8").tofile(ip)
        env = dict(os.environ, OMP_NUM_THREADS="4")
        subprocess.run([HARNESS, jp, ip, op], check=True, env=env, stdout=subprocess.DEVNULL)
        out = np.fromfile(op, dtype="<f8")
    assert out.size == n_out, (out.size, n_out)
    return out


def cfg_json(cfg):
    return np.array(json.dumps(cfg.as_dict()))


# -------------------------------------------------------------------------- fixtures
def make_weno():
    rng = np.random.default_rng(1001)
    parts = []
    x = np.linspace(-2, 2, 6)
    for _ in range(400):  # smooth polynomials / sines at several scales
        a = rng.normal(size=4)
        scale = 10.0 ** rng.uniform(-8, 3)
        parts.append(scale * (a[0] + a[1] * x + a[2] * x ** 2 + a[3] * np.sin(x)))
    for _ in range(400):  # jumps
        lo, hi = rng.uniform(-3, 3, size=2)
        cut = rng.integers(1, 6)
        parts.append(np.where(np.arange(6) < cut, lo, hi))
    for _ in range(200):  # noise
        parts.append(rng.normal(size=6) * 10.0 ** rng.uniform(-3, 3))
    parts.append(np.zeros(6))
    parts.append(np.full(6, 1.7))
    parts.append(np.full(6, -3.25e-7))
    parts.append(np.array([0, 0, 0, 1, 1, 1.0]))
    stencils = np.array(parts, dtype=np.float64)
    out = run(job_text("weno", n=len(stencils)), [stencils], 4 * len(stencils)).reshape(-1, 4)
    return {"stencils": stencils, "out": out}


def physical_pairs(rng, shape, strong=True):
    """Random primitive states -> conserved, for Riemann-solver fixtures."""
    rho = rng.uniform(0.1, 5.0, size=shape)
    u = rng.uniform(-3.0, 3.0, size=shape)
    v = rng.uniform(-3.0, 3.0, size=shape)
    p = rng.uniform(0.05, 10.0, size=shape)
    return rho, u, v, p


def make_flux():
    out = {}
    nx, ny = 12, 10
    R, C = nx + 8, ny + 8
    rng = np.random.default_rng(2002)
    for scheme in (1, 2, 3, 10):
        for gamma in (1.4, 5.0 / 3.0):
            cfg = make_config(nx, ny, scheme=scheme, gamma=gamma)
            rl, ul, vl, pl = physical_pairs(rng, (R, C))
            Ls = cases.prim_to_cons(rl, ul, vl, pl, gamma)
            # half of the faces get a nearby right state, half an independent one
            rr, ur, vr, pr = physical_pairs(rng, (R, C))
            near = rng.uniform(size=(R, C)) < 0.5
            jit = lambda a: a * (1.0 + 0.01 * rng.uniform(-1, 1, size=(R, C)))
            rr = np.where(near, jit(rl), rr)
            ur = np.where(near, ul + 0.01 * rng.uniform(-1, 1, size=(R, C)), ur)
            vr = np.where(near, vl + 0.01 * rng.uniform(-1, 1, size=(R, C)), vr)
            pr = np.where(near, jit(pl), pr)
            Rs = cases.prim_to_cons(rr, ur, vr, pr, gamma)
            n = 4 * R * C
            res = run(job_text("flux", cfg), [Ls, Rs], 2 * n)
            key = f"s{scheme}_g{'53' if gamma > 1.5 else '14'}"
            out[key + "_cfg"] = cfg_json(cfg)
            out[key + "_L"] = Ls
            out[key + "_R"] = Rs
            out[key + "_F1"] = res[:n].reshape(4, R, C)
            out[key + "_F2"] = res[n:].reshape(4, R, C)
    return out


def hc_face_pairs(rng, R, C, gamma):
    """Face states for the heat-capacity solvers: physical pairs plus the faces that
    reach their special branches -- supersonic pairs (all-left / all-right fluxes),
    identical and resting states (s* on the branch boundary), near-vacuum densities,
    negative pressure (NaN sound speed: the x/y HLLHC fallback asymmetry and HLLCHC's
    right-flux default) and negative or zero density (RoeHC's skipped face)."""
    shape = (R, C)
    rl, ul, vl, pl = physical_pairs(rng, shape)
    rr, ur, vr, pr = physical_pairs(rng, shape)
    kind = rng.choice(8, size=shape, p=[0.40, 0.12, 0.12, 0.08, 0.08, 0.08, 0.07, 0.05])
    near = kind == 1
    jit = lambda a: a * (1.0 + 0.01 * rng.uniform(-1, 1, size=shape))
    rr = np.where(near, jit(rl), rr)
    ur = np.where(near, ul + 0.01 * rng.uniform(-1, 1, size=shape), ur)
    vr = np.where(near, vl + 0.01 * rng.uniform(-1, 1, size=shape), vr)
    pr = np.where(near, jit(pl), pr)
    # strong, supersonic pairs: large ratios and speeds of either sign
    strong = kind == 2
    rr = np.where(strong, rl * 10.0 ** rng.uniform(-3, 3, size=shape), rr)
    pr = np.where(strong, pl * 10.0 ** rng.uniform(-5, 5, size=shape), pr)
    drift = rng.choice([-1.0, 1.0], size=shape) * rng.uniform(5, 40, size=shape)
    ul = np.where(strong, ul + drift, ul)
    ur = np.where(strong, ur + drift, ur)
    vl = np.where(strong, vl + drift, vl)
    vr = np.where(strong, vr + drift, vr)
    # identical states; resting states with a pressure jump
    same = kind == 3
    rr, ur, vr, pr = (np.where(same, a, b) for a, b in ((rl, rr), (ul, ur), (vl, vr), (pl, pr)))
    rest = kind == 4
    ul, ur, vl, vr = (np.where(rest, 0.0, a) for a in (ul, ur, vl, vr))
    # near vacuum on one side
    vac = kind == 5
    rl = np.where(vac & (rng.uniform(size=shape) < 0.5), 1e-10, rl)
    rr = np.where(vac, np.where(rl == 1e-10, rr, 1e-9), rr)
    # negative pressure on one or both sides
    negp = kind == 6
    pl = np.where(negp & (rng.uniform(size=shape) < 0.7), -rng.uniform(1e-3, 1.0, size=shape), pl)
    pr = np.where(negp & (rng.uniform(size=shape) < 0.7), -rng.uniform(1e-3, 1.0, size=shape), pr)
    Ls = cases.prim_to_cons(rl, ul, vl, pl, gamma)
    Rs = cases.prim_to_cons(rr, ur, vr, pr, gamma)
    # negative / zero density in the conserved state (momentum and energy kept)
    negr = kind == 7
    pick = rng.uniform(size=shape)
    Ls[0] = np.where(negr & (pick < 0.4), -Ls[0], Ls[0])
    Rs[0] = np.where(negr & (pick >= 0.4) & (pick < 0.8), -Rs[0], Rs[0])
    Ls[0] = np.where(negr & (pick >= 0.8), 0.0, Ls[0])
    return np.ascontiguousarray(Ls), np.ascontiguousarray(Rs)


def make_flux_hc():
    out = {}
    nx, ny = 24, 20
    R, C = nx + 8, ny + 8
    rng = np.random.default_rng(2112)
    for scheme in (11, 22, 33):
        for gamma in (1.4, 5.0 / 3.0):
            cfg = make_config(nx, ny, scheme=scheme, gamma=gamma)
            Ls, Rs = hc_face_pairs(rng, R, C, gamma)
            n = 4 * R * C
            res = run(job_text("flux", cfg), [Ls, Rs], 2 * n)
            key = f"s{scheme}_g{'53' if gamma > 1.5 else '14'}"
            out[key + "_cfg"] = cfg_json(cfg)
            out[key + "_L"] = Ls
            out[key + "_R"] = Rs
            out[key + "_F1"] = res[:n].reshape(4, R, C)
            out[key + "_F2"] = res[n:].reshape(4, R, C)
    # whole steps through the reference's solver switch (include/CFD_convection.h:99-107)
    specs = []
    for scheme in (11, 22, 33):
        for char in (1, 0):
            specs.append((f"ll_s{scheme}_c{char}",
                          lambda s=scheme, c=char: ll_small(20, 16, scheme=s, characteristic=c), 5))
        specs.append((f"sedov_s{scheme}", lambda s=scheme: cases.sedov(18, scheme=s), 4))
    specs.append(("wave_s33", lambda: cases.smooth_wave(16, 12, scheme=33), 4))
    for key, builder, steps in specs:
        cfg, U = builder()
        res = run(job_text("step", cfg, steps=steps), [U], steps + U.size)
        out["step_" + key + "_cfg"] = cfg_json(cfg)
        out["step_" + key + "_U0"] = U
        out["step_" + key + "_dts"] = res[:steps]
        out["step_" + key + "_U"] = res[steps:].reshape(cfg.shape)
    return out


def blocky_state(nx, ny, seed, gamma=1.4):
    """Smooth random state with a superposed oblique discontinuity."""
    U = cases.random_smooth_state(nx, ny, seed, gamma=gamma)
    rho, u, v, p = cases.cons_to_prim(U, gamma)
    X, Y = cases.cell_centres(nx, ny, 1.0 / nx, 1.0 / ny)
    jump = X + 0.7 * Y > 0.8
    rho = np.where(jump, rho * 2.5, rho)
    p = np.where(jump, p * 3.0, p)
    return cases.prim_to_cons(rho, u, v, p, gamma)


def make_eigen():
    nx, ny = 10, 8
    cfg = make_config(nx, ny, gamma=1.4)
    U = blocky_state(nx, ny, 3003)
    n = 16 * cfg.R * cfg.C
    res = run(job_text("eigen", cfg), [U], 4 * n)
    shp = (4, 4, cfg.R, cfg.C)
    return {"cfg": cfg_json(cfg), "U": U, "ELx": res[:n].reshape(shp), "ERx": res[n:2 * n].reshape(shp),
            "ELy": res[2 * n:3 * n].reshape(shp), "ERy": res[3 * n:].reshape(shp)}


def make_recon():
    out = {}
    nx, ny = 14, 12
    for order in (1, 2, 3, 5):
        for char in (0, 1):
            cfg = make_config(nx, ny, recon_order=order, characteristic=char)
            U = blocky_state(nx, ny, 4000 + 10 * order + char)
            n = 4 * cfg.R * cfg.C
            res = run(job_text("recon", cfg), [U], 4 * n).reshape(4, 4, cfg.R, cfg.C)
            key = f"r{order}c{char}"
            out[key + "_cfg"] = cfg_json(cfg)
            out[key + "_U"] = U
            out[key + "_CL1"], out[key + "_CR1"], out[key + "_CL2"], out[key + "_CR2"] = res
    return out


def make_bc():
    out = {}
    combos = {
        "outflow": ((0, 0, 0, 0), 9, 7),
        "reflection": ((1, 1, 1, 1), 9, 7),
        "periodic": ((2, 2, 2, 2), 9, 7),
        "inflow": ((3, 3, 3, 3), 9, 7),
        "fixed": ((4, 4, 4, 4), 9, 7),
        "noh": ((1, 0, 1, 0), 8, 8),
        "dmr": ((4, 0, 1, 4), 10, 6),
        "mixed": ((2, 3, 0, 1), 7, 9),
        "tiny_periodic": ((2, 2, 2, 2), 2, 3),
        "tiny_reflection": ((1, 1, 1, 1), 3, 1),
        "tiny_outflow": ((0, 0, 0, 0), 1, 2),
    }
    rng = np.random.default_rng(5005)
    for name, (bc, nx, ny) in combos.items():
        inflow = rng.uniform(0.5, 2.0, size=4)
        fixed = rng.uniform(0.5, 2.0, size=(4, 4))
        cfg = make_config(nx, ny, bc=bc, inflow_state=inflow, fixed_state=fixed)
        U = rng.uniform(-1.0, 2.0, size=cfg.shape)
        res = run(job_text("bc", cfg), [U], U.size).reshape(cfg.shape)
        out[name + "_cfg"] = cfg_json(cfg)
        out[name + "_U"] = U
        out[name + "_out"] = res
    return out


def ll_small(nx, ny, **kw):
    return cases.four_quadrant(nx, ny, cases.LAX_LIU_3, x0=0.8, y0=0.8, perturb=1e-3,
                               seed=nx * 131 + ny, **kw)


def make_residual():
    out = {}
    nx, ny = 16, 12
    specs = []
    for scheme in (1, 2, 3, 10):
        for char in (1, 0):
            specs.append((f"s{scheme}c{char}", dict(scheme=scheme, characteristic=char)))
    specs.append(("src", dict(scheme=2, characteristic=1, source=True, gravity=-1.0,
                              bc=(1, 1, 1, 1))))
    for key, kw in specs:
        cfg, U = ll_small(nx, ny, **kw)
        n = U.size
        res = run(job_text("residual", cfg), [U], 4 * n).reshape(4, 4, cfg.R, cfg.C)
        out[key + "_cfg"] = cfg_json(cfg)
        out[key + "_U0"] = U
        out[key + "_Ubc"], out[key + "_L"], out[key + "_F"], out[key + "_G"] = res
    return out


def make_steps():
    out = {}
    specs = []
    # solver x reconstruction-mode matrix on a perturbed Lax-Liu configuration
    for scheme in (2, 1, 3, 10):
        for char in (1, 0):
            specs.append((f"ll_s{scheme}_c{char}",
                          lambda s=scheme, c=char: ll_small(20, 16, scheme=s, characteristic=c), 5))
    # time orders
    for order in (1, 2):
        specs.append((f"ll_rk{order}",
                      lambda o=order: ll_small(20, 16, time_order=o, characteristic=1), 5))
    # other reconstruction orders, both modes
    for rec in (1, 2, 3):
        for char in (1, 0):
            specs.append((f"ll_r{rec}_c{char}",
                          lambda r=rec, c=char: ll_small(20, 16, recon_order=r, characteristic=c), 5))
    specs.append(("noh", lambda: cases.noh(16), 6))
    specs.append(("noh_roe", lambda: cases.noh(16, scheme=3), 4))
    specs.append(("sedov", lambda: cases.sedov(18), 5))
    specs.append(("wave_periodic", lambda: cases.smooth_wave(16, 12, scheme=3), 5))
    specs.append(("gravity", lambda: cases.four_quadrant(
        12, 20, {"NE": (2, 0, 0, 2.5), "NW": (2, 0, 0, 2.5), "SW": (1, 0, 0, 2.5), "SE": (1, 0, 0, 2.5)},
        perturb=1e-2, seed=7, source=True, gravity=-1.0, bc=(1, 1, 1, 1), characteristic=1), 5))
    specs.append(("dmr", lambda: cases.double_mach(24, 8, characteristic=1), 4))
    specs.append(("dmr_hllc", lambda: cases.double_mach(24, 8, scheme=2, characteristic=0), 4))
    specs.append(("inflow", lambda: cases.four_quadrant(
        16, 10, cases.LAX_LIU_3, perturb=1e-3, seed=3, bc=(3, 0, 2, 2)), 4))
    specs.append(("fixed_dt", lambda: ll_small(12, 12, characteristic=1), 3))
    for key, builder, steps in specs:
        cfg, U = builder()
        extra = {"steps": steps}
        if key == "fixed_dt":
            extra["fixed_dt"] = 1.0e-3
        res = run(job_text("step", cfg, **extra), [U], steps + U.size)
        out[key + "_cfg"] = cfg_json(cfg)
        out[key + "_U0"] = U
        out[key + "_dts"] = res[:steps]
        out[key + "_U"] = res[steps:].reshape(cfg.shape)
        if "fixed_dt" in extra:
            out[key + "_fixed_dt"] = np.array(extra["fixed_dt"])
    return out


def make_dt():
    out = {}
    for key, (cfg, U) in {"ll": ll_small(20, 16), "noh": cases.noh(16), "sedov": cases.sedov(18),
                          "dmr": cases.double_mach(24, 8)}.items():
        res = run(job_text("dt", cfg), [U], 1)
        out[key + "_cfg"] = cfg_json(cfg)
        out[key + "_U"] = U
        out[key + "_dt"] = res
    return out


def dmr_reference_config(nx, ny):
    """The constants BC_DoubleMach_2D hard-codes (include/Boundary_Condition.h:483-512):
    [0,4]x[0,1], gamma 1.4, pre-shock (1.4, 0, 0, 1), post-shock (8, 7.1447, -4.125,
    116.5), wall and shock start at x = 1/6, slope sqrt(3), shock speed 20 -- as the
    double-Mach-reflection sides of this repository's configuration."""
    import math
    post = tuple(float(x) for x in cases.prim_to_cons(8.0, 7.1447, -4.125, 116.5, 1.4))
    pre = tuple(float(x) for x in cases.prim_to_cons(1.4, 0.0, 0.0, 1.0, 1.4))
    cfg = make_config(nx, ny, dx=4.0 / nx, dy=1.0 / ny, bc=(3, 0, 5, 6), inflow_state=post,
                      fixed_state=(post, post, post, post))
    cfg.lx = 4.0
    cfg.bc_par[2][0] = 1.0 / 6.0
    for k, v in enumerate((1.0 / 6.0, 1.0, math.sqrt(3.0), 20.0)):
        cfg.bc_par[3][k] = v
    for k in range(4):
        cfg.bc_alt_state[3][k] = pre[k]
    return cfg


def make_dmrbc():
    """BC_DoubleMach_2D on random states at several times (the shock trace sweeps the
    top boundary; sizes put cell centres on both sides of the wall start)."""
    out = {}
    rng = np.random.default_rng(7007)
    for name, (nx, ny, t) in {"t0": (48, 12, 0.0), "t1": (48, 12, 0.013), "t2": (60, 9, 0.2),
                              "t3": (24, 20, 0.11)}.items():
        cfg = dmr_reference_config(nx, ny)
        U = rng.uniform(0.5, 2.0, size=cfg.shape)
        res = run(job_text("dmrbc", cfg, time=t), [U], U.size).reshape(cfg.shape)
        out[name + "_cfg"] = cfg_json(cfg)
        out[name + "_t"] = np.array(t)
        out[name + "_U"] = U
        out[name + "_out"] = res
    return out


def state_digest(U):
    """SHA-256 of a state's bytes with -0.0 read as +0.0 (the reference and the device
    may differ in the sign of a zero momentum, which no comparison here counts)."""
    import hashlib
    A = np.ascontiguousarray(np.where(np.asarray(U) == 0.0, 0.0, U), dtype="<f8")
    return hashlib.sha256(A.tobytes()).hexdigest()


def make_fulllength():
    """Whole runs of the reference's own time loop (src/RP_FVM.c:179-191, mode "run"):
    Sedov 200^2 to t = 1.0 (WENO5-char + HLLC, RK3, CFL 0.5, outflow) and Lax-Liu 3 with
    the interface at 0.8 on 512^2 to t = 0.3.  Stored: the input state, every dt, the
    final time, the SHA-256 of the final array (ghosts included), and per-variable
    interior sums and a sparse sample of the final interior for diagnosis."""
    from oracle.pyoracle import harness_run
    out = {}
    for name, (cfg, U0), t_end, stride in (("sedov200", cases.sedov(200), 1.0, 8),
                                           ("ll3_512", cases.lax_liu_3(512), 0.3, 16)):
        dts, t, U, rec = harness_run(cfg, U0, t_end, threads=8)
        print(f"{name}: {len(dts)} steps to t={t} in {rec['seconds']:.1f} s")
        inner = U[:, 4:-4, 4:-4]
        out[f"{name}_cfg"] = cfg_json(cfg)
        out[f"{name}_U0"] = U0
        out[f"{name}_t_end"] = np.array(t_end)
        out[f"{name}_dts"] = dts
        out[f"{name}_t"] = np.array(t)
        out[f"{name}_digest"] = np.array(state_digest(U))
        out[f"{name}_sums"] = inner.reshape(4, -1).sum(axis=1)
        out[f"{name}_stride"] = np.array(stride)
        out[f"{name}_sample"] = inner[:, ::stride, ::stride].copy()
    return out


def main():
    if not os.path.exists(HARNESS):
        sys.exit(f"{HARNESS} missing: run `make -C oracle ref` first")
    makers = (("weno", make_weno), ("flux", make_flux), ("eigen", make_eigen),
              ("recon", make_recon), ("bc", make_bc), ("residual", make_residual),
              ("steps", make_steps), ("dt", make_dt), ("dmrbc", make_dmrbc),
              ("flux_hc", make_flux_hc), ("fulllength", make_fulllength))
    only = set(sys.argv[1:])  # fixture names to (re)write; all when empty
    for name, fn in makers:
        if only and name not in only:
            continue
        data = fn()
        path = os.path.join(OUT, f"{name}.npz")
        np.savez_compressed(path, **data)
        print(f"{path}: {len(data)} arrays, {os.path.getsize(path) / 1024:.1f} KiB")


if __name__ == "__main__":
    main()
