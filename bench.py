#!/usr/bin/env python3
"""Throughput of the 2-D Euler finite-volume step on one MI355X.

Workload (the headline configuration): 2-D Noh implosion, 4096 x 4096 cells on
[0,1]^2, gamma = 5/3, rho = 1, p = 1e-6, unit inflow speed toward the origin,
reflecting walls at x = 0 and y = 0 and outflow at the far walls; characteristic
WENO5 (WENO5-Z) reconstruction, HLLC flux, SSP-RK3, CFL 0.4.  One "step" is one full
RK3 step (dt reduction + 3 stages) over the grid; the unit is Mcell-updates/s
(interior cells x steps / s), inputs resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line.  The CPU baseline is the reference's own FP64 OpenMP step
(oracle/_ref/ref_harness) on a bounded sample of the same workload, on this host's
cores; without that binary it falls back to the oracle port.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mcell-updates/s WENO5-HLLC 4096^2: mixed vs FP64 GPU at L1(rho)<=1e-5; %HBM"
HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes per cell-update of an SSP-RK3 step with FP64 state (4 vars):
# stage 1 reads U^n and writes U1 (64 B), stages 2 and 3 read the stage input and
# U^n and write the output (96 B each); the dt reduction is folded in.
STEP_BYTES_PER_CELL = 256.0
# Peak VALU issue of the card: 256 CUs x 4 SIMDs x 2.4 GHz, one wave64 FP32/FP64 VALU
# instruction per SIMD every 4 cycles (16 lanes per pass).
VALU_PEAK_WAVE_INSTR_PER_S = 256 * 4 * 2.4e9 / 4
# Kernel names behind each timing class, per form (the y sweep is fused with the
# stage update whenever the stage writes a separate output buffer).
KERNELS = {"fp64": {"flux_x": "k_flux_ring", "flux_y": "k_yflux_march", "update": "k_update",
                    "dt": "k_dt_slots", "bc": "k_bc_side"},
           "mixed": {"flux_x": "k_flux_mixed", "flux_y": "k_yflux_march_mixed",
                     "update": "k_update", "dt": "k_dt_slots", "bc": "k_bc_side"}}


def full_length_record(n):
    """The whole-run accuracy record of the workload (profiles/, written by --noh-full):
    mixed vs FP64 L1(rho) at t = 0.6, next to FP64 started from a 1e-12 perturbation, over
    the whole grid and with the origin corner, wall bands and shock band masked."""
    from overheatingproblem_amd.profiling import load_traffic
    rec = load_traffic(f"noh{n}").get("accuracy_full")
    if not rec:
        return None
    return dict(rec, source=f"profiles/noh{n}_t0.6_accuracy.md (python bench.py --noh-full)")


def one_pass_record(prof, rop, rmx, cells, steps, alg_bytes, value):
    """The mixed form's one-pass stage (x faces, y faces and update in one kernel, forced
    with RPFVM_DEBUG_ONE_PASS) beside the default two-sweep stage: throughput and, from
    the committed rocprofv3 counters, the HBM bytes one stage moves."""
    v = cells * steps / (rop["gpu_ms"] / 1e3) / 1e6
    k = prof.get("kernels", {})
    rec = {"value": round(v, 3), "ms_per_step": round(rop["gpu_ms"] / steps, 4),
           "vs_two_sweep": round(v / value, 4), "faults": rop["faults"],
           "same_state_as_two_sweep": bool(np.array_equal(rop["U"], rmx["U"])),
           "kernel": "k_stage_mixed"}
    if "k_stage_mixed" in k:
        tb = k["k_stage_mixed"]["traffic_bytes"]
        rec.update({"traffic": round(tb), "traffic_vs_alg": round(tb / alg_bytes, 3),
                    "two_sweep_stage_traffic_vs_alg": round(
                        (k["k_flux_mixed"]["traffic_bytes"] + k["k_yflux_march_mixed"]["traffic_bytes"])
                        / alg_bytes, 3) if "k_flux_mixed" in k and "k_yflux_march_mixed" in k else None})
    return rec


def kernel_io_bytes(cls, nx, ny, stores_g=True):
    """Bytes one launch of a timing class must read and write (FP64, 4 vars), averaged
    over the three RK3 stages.  x sweep: the stage state and the x-face fluxes.  Fused
    y sweep + update: the stage state, the x-face fluxes, U^n (stages 2 and 3), the
    stage output and, where the form stores them (FP64), the y-face fluxes; the mixed
    form stores only the few its fix-up reads."""
    cell, fx, fy = 32.0 * nx * ny, 32.0 * (nx + 1) * ny, 32.0 * nx * (ny + 1)
    if cls == "flux_x":
        return cell + fx
    if cls == "flux_y":
        return cell + fx + (fy if stores_g else 0.0) + cell + cell * 2.0 / 3.0
    if cls == "dt":
        return cell
    return 2.0 * cell


def log(msg):
    print(msg, file=sys.stderr, flush=True)


class HipEvents:
    """Minimal hipEvent_t binding (timing on a given stream)."""

    def __init__(self):
        self.h = ctypes.CDLL("libamdhip64.so")
        self.h.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.h.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.h.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p,
                                               ctypes.c_void_p]
        self.h.hipEventDestroy.argtypes = [ctypes.c_void_p]

    def create(self):
        e = ctypes.c_void_p()
        assert self.h.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.h.hipEventRecord(e, ctypes.c_void_p(stream)) == 0

    def elapsed_ms(self, a, b):
        assert self.h.hipEventSynchronize(b) == 0
        f = ctypes.c_float()
        assert self.h.hipEventElapsedTime(ctypes.byref(f), a, b) == 0
        return float(f.value)


def host_cpus():
    """CPUs this process may run on (its affinity mask)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def host_cores():
    """Threads of the CPU baseline: every CPU of the affinity mask, or OMP_NUM_THREADS
    when the environment sets it lower (a shared GPU host sets it to the CPU share that
    goes with one GPU)."""
    n = host_cpus()
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_info(cores):
    """What the CPU figures ran on."""
    env = os.environ.get("OMP_NUM_THREADS")
    return {"cpu_model": cpu_model(), "cpus_visible": host_cpus(),
            "omp_num_threads_env": env,
            "why": ("OMP_NUM_THREADS of the environment (the host's CPU share for this GPU)"
                    if env and env.isdigit() and int(env) < host_cpus()
                    else "every CPU of the affinity mask")}


def run_gpu(cfg, U0, steps, warmup, dts=None, timing=True):
    """Warm-up + timed steps through the device-side loop (rpfvm_advance_to: dt and clock on
    the device, no host round trip per step), bracketed by HIP events on the context
    stream, with the library's per-kernel events on (`timing`).  With `dts`: the
    host-driven loop instead (rpfvm_time_advance with the given dts: the replay).  Returns
    timings, kernel times, fault bits and the final state."""
    from overheatingproblem_amd.capi import Solver
    ev = HipEvents()
    with Solver(cfg) as s:
        s.set_state(U0)
        tw = 0.0
        if dts is not None:
            for dt in dts[:warmup]:
                s.time_advance(dt)
        else:
            tw, _ = s.advance_to(float("inf"), warmup)
        faults = s.sync()
        stream = s.stream()
        s.set_kernel_timing(timing)
        s.kernel_times()  # reset
        e0, e1 = ev.create(), ev.create()
        ev.record(e0, stream)
        t0 = time.perf_counter()
        if dts is not None:
            for dt in dts[warmup:warmup + steps]:
                s.time_advance(dt)
            n = len(dts[warmup:warmup + steps])
        else:  # from the warm-up's end: the moving boundary sides read the clock
            _, n = s.advance_to(float("inf"), steps, tw)
        ev.record(e1, stream)
        gpu_ms = ev.elapsed_ms(e0, e1)
        t1 = time.perf_counter()
        kt = s.kernel_times()
        faults |= s.sync()
        U = s.get_state()
    assert n == steps
    return {"gpu_ms": gpu_ms, "wall_s": t1 - t0, "kernels": kt, "faults": faults, "U": U}


def host_loop_dts(cfg, U0, steps):
    """The host-driven loop (rpfvm_get_delta_t, then rpfvm_time_advance), untimed: the dts
    and the final state, for the mixed form's replay and as a check of the device loop."""
    from overheatingproblem_amd.capi import Solver
    dts = []
    with Solver(cfg) as s:
        s.set_state(U0)
        for _ in range(steps):
            dts.append(s.get_delta_t())
            s.time_advance(dts[-1])
        U = s.get_state()
        s.sync()
    return dts, U


def cpu_baseline(n, steps, cores):
    """Reference FP64 OpenMP step (or the oracle port) on `steps` steps of the same
    workload at n^2."""
    from overheatingproblem_amd import cases
    cfg, U = cases.noh(n)
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="true")
    if os.path.exists(harness):
        from oracle.pyoracle import harness_job
        with tempfile.TemporaryDirectory() as td:
            with open(os.path.join(td, "job.txt"), "w") as f:
                f.write(harness_job("bench", cfg, steps=steps))
            U.astype("<f8").tofile(os.path.join(td, "in.bin"))
            r = subprocess.run([harness, os.path.join(td, "job.txt"), os.path.join(td, "in.bin"),
                                os.path.join(td, "out.bin")], capture_output=True, text=True,
                               env=env, check=True)
        rec = json.loads(r.stdout.strip().splitlines()[-1])
        secs, kind, threads = rec["seconds"], "reference", rec["threads"]
    else:
        os.environ["OMP_NUM_THREADS"] = str(cores)
        from oracle import pyoracle
        t0 = time.perf_counter()
        for _ in range(steps):
            pyoracle.time_advance(cfg, U, pyoracle.delta_t(cfg, U))
        secs, kind, threads = time.perf_counter() - t0, "port", pyoracle.threads()
    value = n * n * steps / secs / 1e6
    return {"value": round(value, 4), "unit": "Mcell-updates/s", "cores": threads, "kind": kind,
            "sample": f"{steps} RK3 steps of the same Noh workload at {n}x{n} "
                      f"(dt + 3 stages, IO excluded), {secs:.2f} s",
            "host": host_info(threads)}


def c1_full(cores):
    """The Sedov-type blast on [-1,1]^2 at 200^2 to t = 1.0 (WENO5-char + HLLC, SSP-RK3,
    CFL 0.5, outflow): the reference's own time loop on this host's cores
    (oracle/_ref/ref_harness, timed as src/RP_FVM.c:181-216: dt + step, IO excluded),
    then the FP64 form on the GPU through the same loop, and whether the two final
    states and dt sequences are bit-identical."""
    from overheatingproblem_amd import cases
    from overheatingproblem_amd.capi import Solver
    from oracle.pyoracle import harness_run
    cfg, U0 = cases.sedov(200)
    t_end = 1.0
    log(f"bench: C1 sedov 200^2 to t={t_end}: reference on {cores} threads ...")
    rdts, rt, rU, rec = harness_run(cfg, U0, t_end, threads=cores)
    log(f"bench: reference {len(rdts)} steps in {rec['seconds']:.2f} s; GPU FP64 ...")
    ev = HipEvents()
    dts = []
    with Solver(cfg) as s:
        s.set_state(U0)
        stream = s.stream()
        e0, e1 = ev.create(), ev.create()
        t0 = time.perf_counter()
        ev.record(e0, stream)
        t = 0.0
        while t < t_end:  # src/RP_FVM.c:179-191
            dt = s.get_delta_t()
            if t + dt > t_end:
                dt = t_end - t
            if t + dt >= t_end:
                dt = t_end - t
            dts.append(dt)
            s.time_advance(dt)
            t = t + dt
        ev.record(e1, stream)
        gpu_ms = ev.elapsed_ms(e0, e1)
        wall = time.perf_counter() - t0
        U = s.get_state()
        faults = s.sync()
    cells = cfg.nx * cfg.ny
    log("bench: GPU FP64 through the device-side loop ...")
    with Solver(cfg) as s:
        s.set_state(U0)
        stream = s.stream()
        e0, e1 = ev.create(), ev.create()
        t0 = time.perf_counter()
        ev.record(e0, stream)
        td, nd = s.advance_to(t_end, 10 ** 7)
        ev.record(e1, stream)
        gpu_ms_d = ev.elapsed_ms(e0, e1)
        wall_d = time.perf_counter() - t0
        Ud = s.get_state()
        faults_d = s.sync()

    def same(a, b):
        a, b = np.asarray(a), np.asarray(b)
        return bool(a.shape == b.shape and np.array_equal(np.where(a == 0, 0.0, a),
                                                          np.where(b == 0, 0.0, b)))
    return {"workload": "sedov-200x200 to t=1.0", "steps": len(rdts),
            "gpu_fp64_device_loop": {"value": round(cells * nd / wall_d / 1e6, 3),
                                     "unit": "Mcell-updates/s", "seconds_wall": round(wall_d, 3),
                                     "gpu_ms": round(gpu_ms_d, 2), "steps": nd, "faults": faults_d,
                                     "t": td, "state_bit_identical": same(Ud, rU),
                                     "note": "rpfvm_advance_to: dt, clipping and clock on the device"},
            "cpu_reference": {"value": round(cells * len(rdts) / rec["seconds"] / 1e6, 4),
                              "unit": "Mcell-updates/s", "seconds": round(rec["seconds"], 3),
                              "cores": rec["threads"], "kind": "reference",
                              "host": host_info(rec["threads"])},
            "gpu_fp64": {"value": round(cells * len(dts) / wall / 1e6, 3),
                         "unit": "Mcell-updates/s", "seconds_wall": round(wall, 3),
                         "gpu_ms": round(gpu_ms, 2), "steps": len(dts), "faults": faults,
                         "note": "host-driven loop: one dt read back per step (launch-bound at 200^2)"},
            "bit_identical": {"dts": same(dts, rdts), "state": same(U, rU), "t": t == rt}}


def run_batch(steps, warmup, precision=1):
    """The batching form on the seeded 64-problem four-quadrant workload (sizes 256..1536,
    quadrant states rho ~ U[0.1,2], p ~ U[0.02,2], u, v ~ U[-1.5,1.5], HLLC/HLL/Roe per
    problem, WENO5-char, RK3, CFL 0.5, positivity safeguard on), FP32 flux + FP64 state
    (precision 1) or FP64: autotune, then time `steps` steps of the whole batch (every
    problem is still active: the smallest needs ~400 steps to reach t = 0.25)."""
    from overheatingproblem_amd import cases
    from overheatingproblem_amd.capi import Batch
    ev = HipEvents()
    cfgs, states = cases.batch_case(64)
    for c in cfgs:
        c.precision = precision
    cells = sum(c.nx * c.ny for c in cfgs)
    with Batch(cfgs) as b:
        for k, U in enumerate(states):
            b.set_state(k, U)
        tuned = b.tune()
        for _ in range(warmup):
            b.step(0.25)
        b.sync()
        s = b.stream()
        e0, e1 = ev.create(), ev.create()
        ev.record(e0, s)
        for _ in range(steps):
            b.step(0.25)
        ev.record(e1, s)
        ms = ev.elapsed_ms(e0, e1)
        t, st, f = b.status()
    active = int((st == warmup + steps).sum())
    value = cells * steps / (ms / 1e3) / 1e6
    return {"value": round(value, 3), "unit": "Mcell-updates/s", "ms_per_step": round(ms / steps, 4),
            "workload": "64 four-quadrant problems, n in {256..1536}, seed 20240611, "
                        "rho U[0.1,2], p U[0.02,2], u,v U[-1.5,1.5], interface (0.5,0.5), "
                        "HLLC/HLL/Roe, t_end 0.25, positivity safeguard",
            "cells": cells, "form": "mixed" if precision else "fp64", "tuning": tuned,
            "problems_active_throughout": active,
            "problems_with_fault_bits": {str(b): int((f == b).sum()) for b in sorted(set(int(x) for x in f))}}


def run_dmr(steps, warmup):
    """Double Mach reflection 8192 x 2048 (moving boundaries, WENO5-char + Roe with the
    HLL(E) positivity fallback, CFL 0.5): FP64 and mixed throughput, and the mixed form's
    L1 / L-inf of rho and p against FP64 (mixed replayed on the FP64 run's dts)."""
    from overheatingproblem_amd import cases
    from overheatingproblem_amd.config import PRECISION_MIXED
    cfg, U0 = cases.dmr(8192, 2048)
    mx = cfg.copy()
    mx.precision = PRECISION_MIXED
    cells = cfg.nx * cfg.ny
    r64 = run_gpu(cfg, U0, steps, warmup, timing=False)
    rmx = run_gpu(mx, U0, steps, warmup, timing=False)
    dts, U64h = host_loop_dts(cfg, U0, warmup + steps)
    rep = run_gpu(mx, U0, steps, warmup, dts=dts, timing=False)
    p64 = cases.cons_to_prim(r64["U"][:, 4:-4, 4:-4], cfg.gamma)
    pmx = cases.cons_to_prim(rep["U"][:, 4:-4, 4:-4], cfg.gamma)

    def rate(r):
        return round(cells * steps / (r["gpu_ms"] / 1e3) / 1e6, 3)

    return {"workload": "dmr-8192x2048", "unit": "Mcell-updates/s", "steps": steps,
            "fp64": rate(r64), "mixed": rate(rmx), "speedup": round(rate(rmx) / rate(r64), 4),
            "l1_rho": float(np.abs(pmx[0] - p64[0]).mean()), "linf_rho": float(np.abs(pmx[0] - p64[0]).max()),
            "l1_p": float(np.abs(pmx[3] - p64[3]).mean()), "linf_p": float(np.abs(pmx[3] - p64[3]).max()),
            "faults": {"fp64": r64["faults"], "mixed": rmx["faults"]},
            "fp64_device_loop_matches_host_driven_loop": bool(np.array_equal(r64["U"], U64h))}


def noh_metrics(states, n, gamma, t_end=0.6):
    """Accuracy figures of Noh states {"fp64": U, "mixed": U} (full arrays with ghosts):
    the wall-heating density deficit 1 - rho/16 at the corner cell and over r < 0.05, the
    post-shock plateau 0.1 < r < 0.15, and mixed-vs-FP64 L1 / L-inf of rho and p."""
    from overheatingproblem_amd import cases
    X, Y = cases.cell_centres(n, n, 1.0 / n, 1.0 / n, ghosts=False)
    r = np.sqrt(X * X + Y * Y)
    out = {}
    prim = {k: cases.cons_to_prim(U[:, 4:-4, 4:-4], gamma) for k, U in states.items()}
    for k, (rho, u, v, p) in prim.items():
        out[k] = {"rho_corner": float(rho[0, 0]),
                  "deficit_corner": float(1.0 - rho[0, 0] / 16.0),
                  "deficit_r_lt_0.05": float(1.0 - rho[r < 0.05].mean() / 16.0),
                  "rho_plateau_0.1_0.15": float(rho[(r > 0.1) & (r < 0.15)].mean()),
                  "shock_radius_exact": t_end / 3.0}
    if "fp64" in prim and "mixed" in prim:
        (r64, _, _, p64), (rmx, _, _, pmx) = prim["fp64"], prim["mixed"]
        out["mixed_vs_fp64"] = {
            "l1_rho": float(np.abs(rmx - r64).mean()), "linf_rho": float(np.abs(rmx - r64).max()),
            "l1_rho_rel": float(np.abs(rmx - r64).mean() / np.abs(r64).mean()),
            "l1_p": float(np.abs(pmx - p64).mean()), "linf_p": float(np.abs(pmx - p64).max()),
            "l1_p_rel": float(np.abs(pmx - p64).mean() / np.abs(p64).mean())}
    return out


def noh_regions(n, t):
    """Masks of the Noh interior (cell centres) where round-off grows: the origin corner
    r < 0.05, bands of 0.02 along the two reflecting walls, and a band of +-0.01 around the
    exact shock radius t/3; "rest" is everything else, split into the post-shock "plateau"
    (r < t/3) and the pre-shock "inflow" (r > t/3)."""
    from overheatingproblem_amd import cases
    X, Y = cases.cell_centres(n, n, 1.0 / n, 1.0 / n, ghosts=False)
    r = np.sqrt(X * X + Y * Y)
    corner = r < 0.05
    axes = ((X < 0.02) | (Y < 0.02)) & ~corner
    shock = (np.abs(r - t / 3.0) < 0.01) & ~corner & ~axes
    rest = ~(corner | axes | shock)
    return {"corner": corner, "axes": axes, "shock": shock, "rest": rest,
            "plateau": rest & (r < t / 3.0), "inflow": rest & (r > t / 3.0)}


def region_l1(a, b, masks):
    """L1 (mean |a - b|) over the whole interior and over each region, each region's share
    of the total, and the interior cell (i, j) of the largest difference."""
    e = np.abs(a - b)
    tot = float(e.sum())
    at = np.unravel_index(int(np.argmax(e)), e.shape)
    out = {"l1": float(e.mean()), "linf": float(e.max()), "linf_at": [int(at[0]), int(at[1])]}
    for k, m in masks.items():
        out[f"l1_{k}"] = float(e[m].mean()) if m.any() else 0.0
        out[f"share_{k}"] = float(e[m].sum() / tot) if tot > 0 else 0.0
    out["l1_masked"] = out["l1_rest"]
    return out


def lockstep_history(cfg, U0, t_end, others, every, label):
    """Run the FP64 form with cfg through the driver loop (src/RP_FVM.c:179-191) and every
    (name, cfg, U0) of `others` on the same dts in lockstep; every `every` steps (and at
    the end) compare each other run's density with the FP64 run's by region.  Returns
    (history, final states, FP64 dts, wall time)."""
    from overheatingproblem_amd.capi import Solver
    solvers = {"fp64": Solver(cfg)}
    solvers["fp64"].set_state(U0)
    for name, c, U in others:
        solvers[name] = Solver(c)
        solvers[name].set_state(U)
    hist, dts, t, steps = [], [], 0.0, 0
    t0 = time.perf_counter()
    try:
        def snapshot():
            U64 = solvers["fp64"].get_state()
            r64 = cases_cons_to_prim(U64, cfg.gamma)[0]
            row = {"step": steps, "t": t}
            masks = label["regions"](t, r64)
            for name in solvers:
                if name == "fp64":
                    continue
                rx = cases_cons_to_prim(solvers[name].get_state(), cfg.gamma)[0]
                row[name] = region_l1(rx, r64, masks)
            hist.append(row)
            log(f"bench: {label['name']} step {steps} t={t:.4f}: " +
                ", ".join(f"{k} L1 {row[k]['l1']:.3e} (rest {row[k]['l1_rest']:.3e})"
                          for k in solvers if k != "fp64"))
        while t < t_end:
            dt = solvers["fp64"].get_delta_t()
            if t + dt > t_end:
                dt = t_end - t
            if t + dt >= t_end:
                dt = t_end - t
            for s in solvers.values():
                s.time_advance(dt)
            dts.append(dt)
            t = t + dt
            steps += 1
            if steps % every == 0:
                snapshot()
        if steps % every:
            snapshot()
        finals = {k: s.get_state() for k, s in solvers.items()}
        faults = {k: s.sync() for k, s in solvers.items()}
    finally:
        for s in solvers.values():
            s.close()
    return hist, finals, faults, dts, time.perf_counter() - t0


def cases_cons_to_prim(U, gamma):
    from overheatingproblem_amd.cases import cons_to_prim
    return cons_to_prim(U[:, 4:-4, 4:-4], gamma)


def run_own_dts(cfg, U0, t_end):
    from overheatingproblem_amd.capi import Solver
    t0 = time.perf_counter()
    with Solver(cfg) as s:
        s.set_state(U0)
        t, steps = s.advance_to(t_end, 10 ** 7)
        return s.get_state(), steps, s.sync(), time.perf_counter() - t0


def noh_full(n, t_end=0.6, every=500):
    """Both forms to t_end on the Noh workload.  FP64 runs the driver loop; FP64 from a
    density perturbed by 1e-12 (relative) and the mixed form run on the FP64 run's dts in
    lockstep, and every `every` steps their densities are compared with FP64's over the
    whole grid and by region (origin corner r < 0.05, wall bands, shock band, rest):
    where round-off grows, and the metric's L1 on the part of the grid where FP64 itself
    is round-off stable.  Then the mixed form once more on its own dts.  Wall-heating
    figures of every final state as :func:`noh_metrics`."""
    from overheatingproblem_amd import cases
    from overheatingproblem_amd.config import PRECISION_MIXED
    cfg, U0 = cases.noh(n, positivity=True)
    mx = cfg.copy()
    mx.precision = PRECISION_MIXED
    rng = np.random.default_rng(12)
    Up = U0.copy()
    Up[0] *= 1.0 + 1e-12 * rng.uniform(-1.0, 1.0, size=Up[0].shape)
    label = {"name": f"noh {n}^2", "regions": lambda t, r64: noh_regions(n, t)}
    hist, finals, faults, dts, wall = lockstep_history(
        cfg, U0, t_end, [("mixed_replay", mx, U0), ("fp64_perturbed", cfg, Up)], every, label)
    log(f"bench: noh {n}^2 lockstep done: {len(dts)} steps, {wall:.1f} s; mixed on its own dts ...")
    Umx, mx_steps, mx_faults, mx_wall = run_own_dts(mx, U0, t_end)
    finals["mixed"] = Umx
    out = {"workload": f"noh-{n}x{n}", "t_end": t_end, "every": every,
           "regions": "corner r<0.05; axes x<0.02 or y<0.02; shock |r-t/3|<0.01; rest (= plateau "
                      "r<t/3 + inflow r>t/3)",
           "history": hist}
    m = noh_metrics({"fp64": finals["fp64"], "mixed": finals["mixed"]}, n, cfg.gamma, t_end)
    out.update(m)
    for other in ("mixed_replay", "fp64_perturbed"):
        mm = noh_metrics({"fp64": finals["fp64"], "mixed": finals[other]}, n, cfg.gamma, t_end)
        out[other] = mm["mixed"]
        out[other + "_vs_fp64"] = mm["mixed_vs_fp64"]
    masks = noh_regions(n, t_end)
    r64 = cases_cons_to_prim(finals["fp64"], cfg.gamma)[0]
    out["final_regions"] = {k: region_l1(cases_cons_to_prim(finals[k], cfg.gamma)[0], r64, masks)
                            for k in ("mixed", "mixed_replay", "fp64_perturbed")}
    out["fp64"].update({"steps": len(dts), "faults": faults["fp64"], "wall_s": round(wall, 2), "t": t_end})
    out["mixed_replay"].update({"steps": len(dts), "faults": faults["mixed_replay"]})
    out["fp64_perturbed"].update({"steps": len(dts), "faults": faults["fp64_perturbed"]})
    out["mixed"].update({"steps": mx_steps, "faults": mx_faults, "wall_s": round(mx_wall, 2), "t": t_end})
    return out


def dmr_regions(rho):
    """The double-Mach-reflection interior split at its discontinuities: cells whose
    density differs from a neighbour's by more than 10% in the FP64 state are the
    'fronts' (shocks, slip line, jet) and the rest is smooth flow."""
    d = np.zeros_like(rho, dtype=bool)
    jx = np.abs(np.diff(rho, axis=0)) > 0.1 * np.minimum(rho[1:], rho[:-1])
    jy = np.abs(np.diff(rho, axis=1)) > 0.1 * np.minimum(rho[:, 1:], rho[:, :-1])
    d[1:] |= jx
    d[:-1] |= jx
    d[:, 1:] |= jy
    d[:, :-1] |= jy
    return {"fronts": d, "rest": ~d}


def dmr_full(nx=8192, ny=2048, t_end=0.2, every=2000):
    """Mach-10 double Mach reflection (moving boundaries, WENO5-char + Roe with the HLL(E)
    positivity fallback, CFL 0.5) to t_end: FP64 through the driver loop; FP64 from a
    density perturbed by 1e-12 (relative) and the mixed form on its dts in lockstep
    (density compared with FP64's every `every` steps, over the whole grid and split into
    fronts and smooth flow); then the mixed form on its own dts (its throughput over the
    whole run)."""
    from overheatingproblem_amd import cases
    from overheatingproblem_amd.config import PRECISION_MIXED
    cfg, U0 = cases.dmr(nx, ny)
    mx = cfg.copy()
    mx.precision = PRECISION_MIXED
    rng = np.random.default_rng(12)
    Up = U0.copy()
    Up[0] *= 1.0 + 1e-12 * rng.uniform(-1.0, 1.0, size=Up[0].shape)
    label = {"name": f"dmr {nx}x{ny}", "regions": lambda t, r64: dmr_regions(r64)}
    hist, finals, faults, dts, wall = lockstep_history(
        cfg, U0, t_end, [("mixed_replay", mx, U0), ("fp64_perturbed", cfg, Up)], every, label)
    log(f"bench: dmr lockstep done: {len(dts)} steps, {wall:.1f} s; mixed on its own dts ...")
    Umx, mx_steps, mx_faults, mx_wall = run_own_dts(mx, U0, t_end)
    g = cfg.gamma
    r64, p64 = (cases_cons_to_prim(finals["fp64"], g)[k] for k in (0, 3))
    masks = dmr_regions(r64)
    cells = nx * ny
    out = {"workload": f"dmr-{nx}x{ny}", "t_end": t_end, "every": every, "history": hist,
           "regions": "fronts: cells whose density differs from a neighbour's by > 10% in the "
                      "FP64 state; rest: smooth flow",
           "fp64": {"steps": len(dts), "faults": faults["fp64"], "lockstep_wall_s": round(wall, 2)},
           "mixed_replay": {"faults": faults["mixed_replay"]},
           "fp64_perturbed": {"faults": faults["fp64_perturbed"]},
           "mixed": {"steps": mx_steps, "faults": mx_faults, "wall_s": round(mx_wall, 2),
                     "mcell_updates_per_s": round(cells * mx_steps / mx_wall / 1e6, 1)},
           "final_regions": {k: region_l1(cases_cons_to_prim(U, g)[0], r64, masks)
                             for k, U in (("mixed_replay", finals["mixed_replay"]),
                                          ("fp64_perturbed", finals["fp64_perturbed"]),
                                          ("mixed", Umx))},
           "final_p": {k: {"l1": float(np.abs(cases_cons_to_prim(U, g)[3] - p64).mean()),
                           "linf": float(np.abs(cases_cons_to_prim(U, g)[3] - p64).max())}
                       for k, U in (("mixed_replay", finals["mixed_replay"]),
                                    ("fp64_perturbed", finals["fp64_perturbed"]), ("mixed", Umx))},
           "fronts_fraction": float(masks["fronts"].mean())}
    return out


def ll3_full(n=2048, t_end=0.3):
    """Lax-Liu configuration 3 (the 2048^2 parity configuration) to t_end in both forms:
    FP64 with its dts recorded, mixed on its own dts and mixed replaying the FP64 dts;
    L1 / L-inf of rho and p of each mixed run against FP64 (absolute and relative to the
    mean |FP64|) -- the accuracy gate of the headline metric, at full length, on a flow
    that does not amplify round-off the way Noh's origin does."""
    from overheatingproblem_amd import cases
    from overheatingproblem_amd.capi import Solver
    from overheatingproblem_amd.config import PRECISION_MIXED
    cfg, U0 = cases.lax_liu_3(n)
    mx = cfg.copy()
    mx.precision = PRECISION_MIXED

    def run(c, dts=None):
        t0 = time.perf_counter()
        used, t, steps = [], 0.0, 0
        with Solver(c) as s:
            s.set_state(U0)
            while (t < t_end) if dts is None else (steps < len(dts)):
                if dts is None:  # the driver loop, src/RP_FVM.c:179-191
                    dt = s.get_delta_t()
                    if t + dt >= t_end:
                        dt = t_end - t
                else:
                    dt = dts[steps]
                used.append(dt)
                s.time_advance(dt)
                t += dt
                steps += 1
            U, faults = s.get_state(), s.sync()
        wall = time.perf_counter() - t0
        log(f"bench: lax-liu-3 {n}^2 to t={t} in {steps} steps, {wall:.1f} s")
        return {"U": U, "t": t, "steps": steps, "faults": faults, "wall_s": round(wall, 2),
                "dts": used}

    from overheatingproblem_amd.cases import cons_to_prim
    r64 = run(cfg)
    res = {"mixed": run(mx), "mixed_replay": run(mx, dts=r64["dts"])}
    p64 = cons_to_prim(r64["U"][:, 4:-4, 4:-4], cfg.gamma)
    out = {"workload": f"lax-liu-3-{n}x{n}", "t_end": t_end,
           "fp64": {"steps": r64["steps"], "t": r64["t"], "faults": r64["faults"],
                    "wall_s": r64["wall_s"]}}
    for k, r in res.items():
        pm = cons_to_prim(r["U"][:, 4:-4, 4:-4], cfg.gamma)
        d = {"steps": r["steps"], "t": r["t"], "faults": r["faults"], "wall_s": r["wall_s"]}
        for name, a, b in (("rho", pm[0], p64[0]), ("p", pm[3], p64[3])):
            e = np.abs(a - b)
            d[f"l1_{name}"] = float(e.mean())
            d[f"l1_{name}_rel"] = float(e.mean() / np.abs(b).mean())
            d[f"linf_{name}"] = float(e.max())
        out[k] = d
    return out


SOLVERS = ((2, "HLLC"), (1, "HLL"), (3, "Roe"), (10, "LLF"), (11, "HLLHC"), (22, "HLLCHC"),
           (33, "RoeHC"))


def noh_solvers(n, t_end=0.6, forms=("fp64", "mixed")):
    """Noh to t_end with the positivity safeguard, once per Riemann solver of the reference
    (schemes 1, 2, 3, 10 and the heat-capacity variants 11, 22, 33, which target exactly
    the wall heating this workload shows) and per form: the density deficit each leaves at
    the wall, the plateau, its fault bits and its Mcell-updates/s, and for the mixed form
    its L1(rho) from the FP64 run of the same solver.  At most 6 n steps (the run needs
    about 4 n), so a solver whose state goes non-finite stops early."""
    from overheatingproblem_amd import cases
    from overheatingproblem_amd.capi import Solver
    from overheatingproblem_amd.config import PRECISION_MIXED
    rows = []
    for scheme, name in SOLVERS:
        cfg0, U0 = cases.noh(n, positivity=True, scheme=scheme)
        states = {}
        for form in forms:
            cfg = cfg0.copy()
            if form == "mixed":
                cfg.precision = PRECISION_MIXED
            with Solver(cfg) as s:
                s.set_state(U0)
                t0 = time.perf_counter()
                t, steps = s.advance_to(t_end, 6 * n)
                faults = s.sync()
                wall = time.perf_counter() - t0
                U = s.get_state()
            finite = bool(np.isfinite(U).all())
            row = {"scheme": scheme, "solver": name, "form": form, "steps": steps, "t": t,
                   "faults": faults, "finite": finite, "wall_s": round(wall, 2),
                   "mcell_updates_per_s": round(n * n * steps / wall / 1e6, 1)}
            if finite:
                row.update(noh_metrics({"fp64": U}, n, cfg.gamma, t_end)["fp64"])
                states[form] = U
            if form == "mixed" and "fp64" in states and finite:
                r64 = cases.cons_to_prim(states["fp64"][:, 4:-4, 4:-4], cfg.gamma)[0]
                rmx = cases.cons_to_prim(U[:, 4:-4, 4:-4], cfg.gamma)[0]
                row["l1_rho_vs_fp64"] = float(np.abs(rmx - r64).mean())
            log(f"bench: noh {n}^2 {name} {form}: t={t:.4f} in {steps} steps, {wall:.1f} s, "
                f"faults {faults}, finite {finite}")
            rows.append(row)
    return {"workload": f"noh-{n}x{n}", "t_end": t_end, "forms": list(forms),
            "positivity_safeguard": True, "solvers": rows}


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=4096, help="grid cells per side")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-dmr", action="store_true")
    ap.add_argument("--noh-full", action="store_true",
                    help="run both forms to t = 0.6 and print the accuracy report instead")
    ap.add_argument("--ll3-full", action="store_true",
                    help="run Lax-Liu 3 at 2048^2 to t = 0.3 in both forms and print the "
                         "mixed-vs-FP64 accuracy instead")
    ap.add_argument("--dmr-full", action="store_true",
                    help="run the double Mach reflection 8192x2048 to t = 0.2 in both forms and "
                         "print the accuracy record instead")
    ap.add_argument("--c1-full", action="store_true",
                    help="run the Sedov 200^2 workload to t = 1.0 with the reference on the host "
                         "cores and the FP64 form on the GPU, and print that record instead")
    ap.add_argument("--noh-solvers", action="store_true",
                    help="run the FP64 form to t = 0.6 once per Riemann solver and print the "
                         "wall-heating report instead")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the library has "
              f"no multi-GPU path; it runs one GPU's share of the step)", file=sys.stderr)
        sys.exit(2)

    if args.noh_full:
        print(json.dumps(noh_full(args.n)), flush=True)
        return
    if args.noh_solvers:
        print(json.dumps(noh_solvers(args.n)), flush=True)
        return
    if args.dmr_full:
        print(json.dumps(dmr_full()), flush=True)
        return
    if args.c1_full:
        print(json.dumps(c1_full(host_cores())), flush=True)
        return
    if args.ll3_full:
        print(json.dumps(ll3_full()), flush=True)
        return
    from overheatingproblem_amd import cases
    from overheatingproblem_amd.config import DEBUG_ONE_PASS, PRECISION_MIXED
    n = args.n
    cfg, U0 = cases.noh(n, positivity=True)
    cfg_mx = cfg.copy()
    cfg_mx.precision = PRECISION_MIXED
    cells = n * n
    log(f"bench: noh {n}x{n} warmup={args.warmup} steps={args.steps}: fp64 ...")
    # the throughput runs record no per-kernel events (each event in the stream costs GPU
    # time); a second timed run of the same steps takes the per-kernel durations
    r64 = run_gpu(cfg, U0, args.steps, args.warmup, timing=False)
    r64["kernels"] = run_gpu(cfg, U0, args.steps, args.warmup)["kernels"]
    log(f"bench: fp64 {r64['gpu_ms'] / args.steps:.3f} ms/step; mixed ...")
    rmx = run_gpu(cfg_mx, U0, args.steps, args.warmup, timing=False)
    rkt = run_gpu(cfg_mx, U0, args.steps, args.warmup)
    rmx["kernels"] = rkt["kernels"]
    log(f"bench: mixed {rmx['gpu_ms'] / args.steps:.3f} ms/step ({rkt['gpu_ms'] / args.steps:.3f} "
        f"with per-kernel events); mixed one-pass stage ...")
    cfg_op = cfg_mx.copy()
    cfg_op.debug_flags = DEBUG_ONE_PASS
    rop = run_gpu(cfg_op, U0, args.steps, args.warmup, timing=False)
    log(f"bench: mixed one-pass {rop['gpu_ms'] / args.steps:.3f} ms/step; host-driven fp64 dts ...")
    dts64, U64h = host_loop_dts(cfg, U0, args.warmup + args.steps)
    log("bench: mixed replay on the fp64 dts (host-driven loop) ...")
    rep = run_gpu(cfg_mx, U0, args.steps, args.warmup, dts=dts64, timing=False)

    def rate(r):
        return cells * args.steps / (r["gpu_ms"] / 1e3) / 1e6

    value, value64 = rate(rmx), rate(r64)
    ms_per_step = rmx["gpu_ms"] / args.steps
    from overheatingproblem_amd.cases import cons_to_prim
    p64 = cons_to_prim(r64["U"][:, 4:-4, 4:-4], cfg.gamma)
    pmx = cons_to_prim(rep["U"][:, 4:-4, 4:-4], cfg.gamma)
    l1 = float(np.abs(pmx[0] - p64[0]).mean())
    linf = float(np.abs(pmx[0] - p64[0]).max())
    l1p = float(np.abs(pmx[3] - p64[3]).mean())
    linfp = float(np.abs(pmx[3] - p64[3]).max())
    log(f"bench: mixed {value:.1f} vs fp64 {value64:.1f} Mcell-updates/s "
        f"(x{value / value64:.2f}), L1(rho) {l1:.3e}")

    kt = rmx["kernels"]
    dom = max(kt, key=lambda k: kt[k][0])
    dom_ms, dom_n = kt[dom]
    avg_ms = dom_ms / max(1, dom_n)
    kname = KERNELS["mixed"].get(dom, dom)
    # Algorithmic bytes of one launch: the step's 256 B per cell-update over its three RK
    # stages (one launch of the fused y sweep completes one stage of every cell).
    alg_bytes = STEP_BYTES_PER_CELL / 3.0 * cells
    achieved = alg_bytes / (avg_ms / 1e3) / 1e9
    io_bytes = kernel_io_bytes(dom, n, n, stores_g=False)  # what the mixed kernel must move
    io_gbs = io_bytes / (avg_ms / 1e3) / 1e9
    step_gbs = STEP_BYTES_PER_CELL * cells / (ms_per_step / 1e3) / 1e9
    from overheatingproblem_amd.profiling import load_traffic
    prof = load_traffic(f"noh{n}")
    tr = prof.get("kernels", {}).get(kname)
    va = prof.get("valu", {}).get(kname)
    # measured HBM bytes of one whole stage of the mixed form (every kernel a stage
    # launches, per launch) against the stage's algorithmic bytes
    stage_tr = None
    kt_all = prof.get("kernels", {})
    if all(k in kt_all for k in ("k_flux_mixed", "k_yflux_march_mixed")):
        per_stage = kt_all["k_flux_mixed"]["traffic_bytes"] + kt_all["k_yflux_march_mixed"]["traffic_bytes"]
        for small in ("k_bc_side", "k_flux_fix", "k_yfix_cells", "k_repair_block"):
            if small in kt_all:  # per stage: 4 ghost sides, 2 fix-ups, 1 repair
                reps = {"k_bc_side": 4, "k_flux_fix": 2}.get(small, 1)
                per_stage += reps * kt_all[small]["traffic_bytes"]
        stage_tr = {"bytes": round(per_stage), "vs_alg": round(per_stage / alg_bytes, 3),
                    "bytes_per_cell_update": round(3.0 * per_stage / cells, 1)}
    valu = None
    if va:
        rate = va["valu_wave_instr_per_dispatch"] / (avg_ms / 1e3)
        valu = {"achieved": round(rate / 1e9, 2), "peak": VALU_PEAK_WAVE_INSTR_PER_S / 1e9,
                "unit": "G wave-instr/s", "frac": round(rate / VALU_PEAK_WAVE_INSTR_PER_S, 4),
                "instr_per_face": round(va["valu_wave_instr_per_dispatch"] * 64 / ((n + 1) * n), 1),
                "source": "SQ_INSTS_VALU per dispatch (profiles/) over this run's launch time"}
    def form(r, v):
        return {"value": round(v, 3), "ms_per_step": round(r["gpu_ms"] / args.steps, 4),
                "step_hbm_frac": round(STEP_BYTES_PER_CELL * cells * v * 1e6 / cells / 8e12, 5),
                "kernel_ms_per_step": {k: round(t[0] / args.steps, 4) for k, t in r["kernels"].items()},
                "faults": r["faults"], "finite": bool(np.isfinite(r["U"]).all())}

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"noh-{n}x{n}", "scheme": "WENO5-char(Z) + HLLC, SSP-RK3",
                   "cfl": 0.4, "gamma": "5/3", "p0": 1e-6, "form": "mixed",
                   "positivity_safeguard": True, "cells": cells},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": round(tr["traffic_bytes"]) if tr else None,
                     "traffic_unit": "B per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
                     "traffic_vs_alg": round(tr["traffic_bytes"] / alg_bytes, 3) if tr else None,
                     "avg_launch_ms": round(avg_ms, 4), "launches": dom_n,
                     "alg_bytes_per_launch": alg_bytes,
                     "alg_basis": "256 B per cell-update / 3 RK stages, one stage per launch",
                     "stage_traffic": stage_tr,
                     "kernel_io": {"bytes_per_launch": io_bytes, "achieved": round(io_gbs, 2),
                                   "frac": round(io_gbs / HBM_PEAK_GBS, 5),
                                   "basis": "the launch's own reads and writes: stage input, "
                                            "x-face fluxes, U^n in 2 of 3 stages, output"}},
        "valu_roofline": valu,
        "step_hbm": {"achieved_gbs": round(step_gbs, 2), "frac": round(step_gbs / HBM_PEAK_GBS, 5),
                     "bytes_per_cell_update": STEP_BYTES_PER_CELL},
        "forms": {"mixed": form(rmx, value), "fp64": form(r64, value64),
                  "speedup_mixed_vs_fp64": round(value / value64, 4)},
        "accuracy": {"l1_rho_mixed_vs_fp64": l1, "linf_rho_mixed_vs_fp64": linf,
                     "l1_p_mixed_vs_fp64": l1p, "linf_p_mixed_vs_fp64": linfp,
                     "note": f"mixed replayed on the fp64 run's {len(dts64)} dts",
                     "full_length": full_length_record(n)},
        "one_pass_stage": one_pass_record(prof, rop, rmx, cells, args.steps, alg_bytes, value),
        "kernel_timing": {"source": "a second timed run of the same steps with the library's "
                                    "per-kernel HIP events on the context stream",
                          "mixed_ms_per_step_with_events": round(rkt["gpu_ms"] / args.steps, 4)},
        "loop": {"kind": "device (rpfvm_advance_to: dt, clipping and clock on the device)",
                 "fp64_matches_host_driven_loop": bool(np.array_equal(r64["U"], U64h)),
                 "host_driven_mixed_ms_per_step": round(rep["gpu_ms"] / args.steps, 4)},
        "wall_s": round(rmx["wall_s"], 4),
    }
    if not args.no_dmr:
        log("bench: double Mach reflection 8192x2048 ...")
        try:
            out["dmr"] = run_dmr(10, 2)
        except Exception as e:  # the headline figure stands without it
            out["dmr"] = {"value": None, "error": str(e)}
    if not args.no_batch:
        log("bench: batch form (64 uneven problems) ...")
        try:
            out["batch"] = {"mixed": run_batch(10, 2, 1), "fp64": run_batch(10, 2, 0)}
        except Exception as e:  # the headline figure stands without it
            out["batch"] = {"value": None, "error": str(e)}
    if not args.no_cpu:
        log(f"bench: cpu baseline ({args.cpu_steps} steps at {n}^2) ...")
        try:
            out["cpu_baseline"] = cpu_baseline(n, args.cpu_steps, host_cores())
        except Exception as e:  # the GPU figure stands without it
            out["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
