"""Summaries of rocprofv3 CSV output for the committed profile records.

    python -m overheatingproblem_amd.profiling trace   <trace_dir> <name> "<command>"
    python -m overheatingproblem_amd.profiling traffic <fetch_dir> <write_dir> \
        <cal_fetch_dir> <cal_write_dir> <cal_bytes> <name> "<commands>"
    python -m overheatingproblem_amd.profiling valu    <pmc_dir> <name> "<command>"
    python -m overheatingproblem_amd.profiling accuracy <noh_full.json> <name> "<command>"
    python -m overheatingproblem_amd.profiling solvers <noh_solvers.json> <name> "<command>"
    python -m overheatingproblem_amd.profiling ll3     <ll3_full.json> <name> "<command>"
    python -m overheatingproblem_amd.profiling dmr     <dmr_full.json> <name> "<command>"
    python -m overheatingproblem_amd.profiling c1      <c1_full.json> <name> "<command>"

`trace` turns a `rocprofv3 --kernel-trace --stats --output-format csv` directory into a
markdown table of per-kernel calls and average durations.  `traffic` turns two
`--pmc FETCH_SIZE` / `--pmc WRITE_SIZE` runs into HBM bytes per dispatch, each counter
scaled by what it reads for a plain copy kernel of `cal_bytes` read and written per
dispatch (overheatingproblem_amd/lib/hbm_probe), and merges the figures into
profiles/hbm_traffic.json, which bench.py reports as roofline.traffic.  `valu` records
per-dispatch SQ instruction counters (SQ_INSTS_VALU ...) in the same file, from which
bench.py reports the VALU issue rate of the dominant kernel.  `accuracy` writes the
full-length Noh accuracy record from the JSON line of `bench.py --noh-full`, `solvers`
the per-solver wall-heating record from `bench.py --noh-solvers`.
"""
from __future__ import annotations

import csv
import glob
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROFILES = os.path.join(ROOT, "profiles")
TRAFFIC_JSON = os.path.join(PROFILES, "hbm_traffic.json")


def _one(d: str, suffix: str) -> str:
    hits = sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True))
    if not hits:
        raise FileNotFoundError(f"no *{suffix} under {d}")
    return hits[0]


def kernel_stats(trace_dir: str) -> list[dict]:
    rows = list(csv.DictReader(open(_one(trace_dir, "kernel_stats.csv"))))
    return [{"name": r["Name"], "calls": int(r["Calls"]), "avg_ns": float(r["AverageNs"]),
             "total_ns": float(r["TotalDurationNs"]), "pct": float(r["Percentage"]),
             "min_ns": float(r["MinNs"]), "max_ns": float(r["MaxNs"])} for r in rows]


def counter_means(pmc_dir: str) -> dict:
    """{kernel: {counter: (mean value per dispatch, dispatches)}}"""
    acc: dict = {}
    for r in csv.DictReader(open(_one(pmc_dir, "counter_collection.csv"))):
        k = r["Kernel_Name"].split("(")[0].strip()
        acc.setdefault(k, {}).setdefault(r["Counter_Name"], {}).setdefault(
            r["Dispatch_Id"], 0.0)
        acc[k][r["Counter_Name"]][r["Dispatch_Id"]] += float(r["Counter_Value"])
    return {k: {c: (sum(v.values()) / len(v), len(v)) for c, v in cs.items()}
            for k, cs in acc.items()}


def trace_markdown(trace_dir: str, name: str, command: str) -> str:
    st = kernel_stats(trace_dir)
    lines = [f"# Kernel trace: {name}", "", "Command (from the repository root):", "",
             "```", command, "```", "",
             "| kernel | calls | avg us | min us | max us | % of GPU time |",
             "|---|---:|---:|---:|---:|---:|"]
    for r in st:
        lines.append(f"| {r['name'][:60]} | {r['calls']} | {r['avg_ns'] / 1e3:.1f} | "
                     f"{r['min_ns'] / 1e3:.1f} | {r['max_ns'] / 1e3:.1f} | {r['pct']:.2f} |")
    return "\n".join(lines) + "\n"


def traffic(fetch_dir, write_dir, cal_fetch_dir, cal_write_dir, cal_bytes: int) -> dict:
    f, w = counter_means(fetch_dir), counter_means(write_dir)
    cf, cw = counter_means(cal_fetch_dir), counter_means(cal_write_dir)
    kib = cal_bytes / 1024.0
    fetch_scale = kib / cf["copy_kernel"]["FETCH_SIZE"][0]
    write_scale = kib / cw["copy_kernel"]["WRITE_SIZE"][0]
    out = {"calibration": {"copy_bytes_per_dispatch": cal_bytes,
                           "fetch_kib_read": cf["copy_kernel"]["FETCH_SIZE"][0],
                           "write_kib_read": cw["copy_kernel"]["WRITE_SIZE"][0],
                           "fetch_scale": fetch_scale, "write_scale": write_scale},
           "kernels": {}}
    for k in sorted(set(f) & set(w)):
        if k.startswith("__amd"):
            continue
        fb = f[k]["FETCH_SIZE"][0] * 1024.0 * fetch_scale
        wb = w[k]["WRITE_SIZE"][0] * 1024.0 * write_scale
        out["kernels"][k] = {"fetch_bytes": fb, "write_bytes": wb, "traffic_bytes": fb + wb,
                             "dispatches": f[k]["FETCH_SIZE"][1]}
    return out


def traffic_markdown(t: dict, name: str, commands: str) -> str:
    c = t["calibration"]
    lines = [f"# HBM traffic per dispatch: {name}", "", "Commands (from the repository root):",
             "", "```", commands, "```", "",
             f"Calibration: the copy kernel reads and writes {c['copy_bytes_per_dispatch']} B "
             f"per dispatch; FETCH_SIZE reads {c['fetch_kib_read']:.0f} KiB (scale "
             f"{c['fetch_scale']:.4f}), WRITE_SIZE reads {c['write_kib_read']:.0f} KiB (scale "
             f"{c['write_scale']:.4f}).  Every figure below is the counter times its scale.", "",
             "| kernel | dispatches | fetch MB | write MB | total MB |",
             "|---|---:|---:|---:|---:|"]
    for k, v in t["kernels"].items():
        lines.append(f"| {k} | {v['dispatches']} | {v['fetch_bytes'] / 1e6:.2f} | "
                     f"{v['write_bytes'] / 1e6:.2f} | {v['traffic_bytes'] / 1e6:.2f} |")
    return "\n".join(lines) + "\n"


def agent(pmc_dir: str) -> dict:
    """CU, SIMD and XCD counts of the GPU agent of a rocprofv3 run."""
    for r in csv.DictReader(open(_one(pmc_dir, "agent_info.csv"))):
        if r.get("Name", "").startswith("gfx"):
            return {"cus": int(r["Cu_Count"]), "simds": int(r["Simd_Count"]),
                    "xcds": int(r["Num_Xcc"]), "name": r["Name"]}
    raise ValueError(f"no GPU agent in {pmc_dir}")


def valu(pmc_dir: str) -> dict:
    """Per-dispatch SQ instruction counts (SQ_INSTS_VALU etc.) of each kernel, and, when
    the run also counted SQ_ACTIVE_INST_VALU, SQ_WAVE_CYCLES and GRBM_GUI_ACTIVE:
      valu_busy       = SQ_ACTIVE_INST_VALU / (CUs x cycles): the fraction of SIMD cycles
                        issuing VALU work (the counter is in quad-cycles, 4 SIMDs per CU);
      waves_per_simd  = 4 SQ_WAVE_CYCLES / (SIMDs x cycles): resident wavefronts;
    with cycles = GRBM_GUI_ACTIVE / XCDs (the counter is summed over the XCDs)."""
    ag = agent(pmc_dir)
    out = {}
    for k, cs in counter_means(pmc_dir).items():
        if k.startswith("__amd") or "SQ_INSTS_VALU" not in cs:
            continue
        rec = {c.lower() + "_per_dispatch": v for c, (v, _) in cs.items()}
        rec["valu_wave_instr_per_dispatch"] = cs["SQ_INSTS_VALU"][0]
        rec["dispatches"] = cs["SQ_INSTS_VALU"][1]
        if all(c in cs for c in ("SQ_ACTIVE_INST_VALU", "SQ_WAVE_CYCLES", "GRBM_GUI_ACTIVE")):
            cycles = cs["GRBM_GUI_ACTIVE"][0] / ag["xcds"]
            if cycles > 0:
                rec["valu_busy"] = cs["SQ_ACTIVE_INST_VALU"][0] / (ag["cus"] * cycles)
                rec["waves_per_simd"] = 4.0 * cs["SQ_WAVE_CYCLES"][0] / (ag["simds"] * cycles)
        out[k] = rec
    return out


def valu_markdown(v: dict, name: str, command: str) -> str:
    cols = sorted({c for rec in v.values() for c in rec if c.endswith("_per_dispatch")})
    derived = [c for c in ("valu_busy", "waves_per_simd") if any(c in r for r in v.values())]
    lines = [f"# SQ instruction counters per dispatch: {name}", "",
             "Command (from the repository root):", "", "```", command, "```", ""]
    if derived:
        lines += ["valu_busy = SQ_ACTIVE_INST_VALU / (CUs x cycles) and waves_per_simd = "
                  "4 SQ_WAVE_CYCLES / (SIMDs x cycles), with cycles = GRBM_GUI_ACTIVE / XCDs "
                  "(SQ_ACTIVE_INST_VALU and SQ_WAVE_CYCLES count quad-cycles; GRBM_GUI_ACTIVE "
                  "is summed over the 8 XCDs).", ""]
    lines += ["| kernel | dispatches | " + " | ".join(derived + [c.replace("_per_dispatch", "")
                                                          for c in cols]) + " |",
              "|---|---:|" + "---:|" * (len(derived) + len(cols))]
    for k, rec in v.items():
        lines.append(f"| {k} | {rec['dispatches']} | "
                     + " | ".join([f"{rec[c]:.3f}" if c in rec else "" for c in derived]
                                  + [f"{rec.get(c, 0):.4g}" for c in cols]) + " |")
    return "\n".join(lines) + "\n"


def _region_row(label: str, r: dict, regions) -> str:
    cells = " | ".join(f"{r['l1_' + k]:.2e} ({r['share_' + k]:.2f})" for k in regions)
    at = f" at {tuple(r['linf_at'])}" if "linf_at" in r else ""
    return f"| {label} | {r['l1']:.3e} | {cells} | {r['linf']:.2f}{at} |"


def _history_table(d: dict, runs, regions) -> list:
    cols = ["l1", "l1_rest"] + (["l1_inflow"] if "inflow" in regions else [])
    names = {"l1": "L1", "l1_rest": "L1 rest", "l1_inflow": "L1 inflow"}
    head = "| step | t | " + " | ".join(f"{lab} {names[c]}" for lab, _ in runs for c in cols) + " |"
    lines = [head, "|" + "---:|" * (2 + len(cols) * len(runs))]
    for h in d["history"]:
        cells = " | ".join(f"{h[k][c]:.2e}" for _, k in runs for c in cols)
        lines.append(f"| {h['step']} | {h['t']:.4f} | {cells} |")
    return lines


def accuracy_markdown(d: dict, command: str) -> str:
    """The Noh full-length accuracy record from `bench.py --noh-full`'s JSON line."""
    runs = (("FP64", "fp64"), ("mixed (own dts)", "mixed"), ("mixed (FP64 dts)", "mixed_replay"),
            ("FP64, rho perturbed 1e-12 (FP64 dts)", "fp64_perturbed"))
    regions = ("corner", "axes", "shock", "rest")
    if "l1_inflow" in d["final_regions"]["mixed"]:
        regions += ("plateau", "inflow")
    n = d["workload"].split("-")[1]
    lines = [f"# Noh {n} to t = {d['t_end']}: accuracy of the mixed form", "",
             "Command (from the repository root):", "", "```", command, "```", "",
             "Noh implosion, gamma 5/3, p0 = 1e-6, WENO5-char + HLLC, SSP-RK3, CFL 0.4, "
             "positivity safeguard on, one MI355X.  Exact solution: rho = 16 behind the shock "
             "(r < t/3); the wall-heating deficit is 1 - rho/16 near the origin.  FP64 runs the "
             "driver loop; the mixed form and FP64 from a density perturbed by 1e-12 "
             "(relative) run on its dts in lockstep, compared with it every "
             f"{d['every']} steps; then the mixed form once more on its own dts.", "",
             "| run | steps | faults | rho(corner cell) | deficit corner | deficit r<0.05 | "
             "rho plateau 0.1<r<0.15 |", "|---|---:|---:|---:|---:|---:|---:|"]
    for label, k in runs:
        r = d[k]
        lines.append(f"| {label} | {r['steps']} | {r['faults']} | {r['rho_corner']:.4f} | "
                     f"{r['deficit_corner']:+.4f} | {r['deficit_r_lt_0.05']:+.5f} | "
                     f"{r['rho_plateau_0.1_0.15']:.4f} |")
    lines += ["", f"Density difference from the FP64 run at t = {d['t_end']}, by region "
              f"({d['regions']}): L1 = mean |difference| over the region's cells, with the "
              "region's share of the whole grid's sum in parentheses; L-inf and its interior "
              "cell (i, j).", "",
              "| compared with FP64 | L1 | " + " | ".join(f"L1 {k}" for k in regions) +
              " | L-inf |", "|---|---:|" + "---:|" * len(regions) + "---|"]
    for label, k in runs[1:]:
        lines.append(_region_row(label, d["final_regions"][k], regions))
    lines += ["", "History (FP64 dts, lockstep): L1 over the whole grid, over the rest "
              "(corner, wall bands and shock band masked) and over its pre-shock inflow part.", ""]
    lines += _history_table(d, (("mixed", "mixed_replay"), ("perturbed FP64", "fp64_perturbed")), regions)
    fr = d["final_regions"]
    ratio = fr["mixed_replay"]["l1_rest"] / fr["fp64_perturbed"]["l1_rest"]
    lines += ["", "Reading: the masked L1 (rest) of the mixed form is "
              f"{fr['mixed_replay']['l1_rest']:.2e} on the FP64 dts, {ratio:.2f} x the distance "
              "FP64 itself travels from a 1e-12 density perturbation "
              f"({fr['fp64_perturbed']['l1_rest']:.2e}); both grow from the origin corner "
              "outwards along the walls and into the post-shock plateau (history)."]
    if "l1_inflow" in fr["mixed_replay"]:
        lines += ["", "In the pre-shock inflow (the part of the rest the shock has not reached) "
                  "the mixed form's L1(rho) is "
                  f"{fr['mixed_replay']['l1_inflow']:.2e} on the FP64 dts and "
                  f"{fr['mixed']['l1_inflow']:.2e} on its own (perturbed FP64: "
                  f"{fr['fp64_perturbed']['l1_inflow']:.2e}); in the post-shock plateau it is "
                  f"{fr['mixed_replay']['l1_plateau']:.2e} (perturbed FP64 "
                  f"{fr['fp64_perturbed']['l1_plateau']:.2e})."]
    return "\n".join(lines) + "\n"


def accuracy_summary(d: dict) -> dict:
    """The figures of the Noh full-length record that bench.py's line carries."""
    fr = d["final_regions"]
    return {"t_end": d["t_end"], "steps_fp64": d["fp64"]["steps"],
            "l1_rho_mixed_own_dts": fr["mixed"]["l1"],
            "l1_rho_mixed_fp64_dts": fr["mixed_replay"]["l1"],
            "l1_rho_fp64_perturbed_1e-12": fr["fp64_perturbed"]["l1"],
            "l1_rho_rest_mixed_fp64_dts": fr["mixed_replay"]["l1_rest"],
            "l1_rho_rest_fp64_perturbed_1e-12": fr["fp64_perturbed"]["l1_rest"],
            "l1_rho_inflow_mixed_own_dts": fr["mixed"].get("l1_inflow"),
            "l1_rho_inflow_mixed_fp64_dts": fr["mixed_replay"].get("l1_inflow"),
            "l1_rho_inflow_fp64_perturbed_1e-12": fr["fp64_perturbed"].get("l1_inflow"),
            "l1_rho_plateau_mixed_fp64_dts": fr["mixed_replay"].get("l1_plateau"),
            "masked": d["regions"]}


def dmr_markdown(d: dict, command: str) -> str:
    """The double-Mach-reflection full-length record from `bench.py --dmr-full`."""
    regions = ("fronts", "rest")
    lines = [f"# Double Mach reflection {d['workload'].split('-')[1]} to t = {d['t_end']}: "
             "accuracy and throughput of the mixed form", "",
             "Command (from the repository root):", "", "```", command, "```", "",
             "Mach-10 double Mach reflection, moving boundaries, WENO5-char + Roe with the "
             "positivity safeguard (HLLE fallback), SSP-RK3, CFL 0.5, one MI355X.  FP64 runs the "
             "driver loop; the mixed form and FP64 from a density perturbed by 1e-12 (relative) "
             f"run on its dts in lockstep, compared every {d['every']} steps; then the mixed "
             f"form on its own dts.  Regions: {d['regions']} ({100 * d['fronts_fraction']:.2f}% "
             "of the cells are fronts).", "",
             "| run | steps | faults | Mcell-updates/s |", "|---|---:|---:|---:|",
             f"| FP64 (driver loop) | {d['fp64']['steps']} | {d['fp64']['faults']} | |",
             f"| mixed (own dts) | {d['mixed']['steps']} | {d['mixed']['faults']} | "
             f"{d['mixed']['mcell_updates_per_s']:.0f} |", "",
             f"Density difference from the FP64 run at t = {d['t_end']} (L1 by region, share in "
             "parentheses), and L1 / L-inf of pressure:", "",
             "| compared with FP64 | L1(rho) | L1 fronts | L1 rest | L-inf(rho) | L1(p) | "
             "L-inf(p) |", "|---|---:|---:|---:|---|---:|---:|"]
    for label, k in (("mixed (FP64 dts)", "mixed_replay"), ("FP64, rho perturbed 1e-12", "fp64_perturbed"),
                     ("mixed (own dts)", "mixed")):
        r, pp = d["final_regions"][k], d["final_p"][k]
        lines.append(f"| {label} | {r['l1']:.3e} | {r['l1_fronts']:.2e} ({r['share_fronts']:.2f}) | "
                     f"{r['l1_rest']:.2e} ({r['share_rest']:.2f}) | {r['linf']:.2f} at "
                     f"{tuple(r['linf_at'])} | {pp['l1']:.3e} | {pp['linf']:.2f} |")
    lines += ["", "History (FP64 dts, lockstep):", ""]
    lines += _history_table(d, (("mixed", "mixed_replay"), ("perturbed FP64", "fp64_perturbed")), regions)
    return "\n".join(lines) + "\n"


def c1_markdown(d: dict, command: str) -> str:
    """The C1 record (Sedov 200^2 to t = 1.0) from `bench.py --c1-full`."""
    c, g, b = d["cpu_reference"], d["gpu_fp64"], d["bit_identical"]
    h = c["host"]
    return "\n".join([
        "# Sedov-type blast 200 x 200 to t = 1.0: the reference's CPU step and the FP64 form", "",
        "Command (from the repository root):", "", "```", command, "```", "",
        "Sedov-type blast on [-1,1]^2, gamma 1.4, rho = 1, p = 1e-5, E0 = 1 in the cells with "
        "r <= 2 dx; WENO5-char + HLLC, SSP-RK3, CFL 0.5, outflow, t_end = 1.0.  The reference's "
        "own step (oracle/_ref/ref_harness, mode run: the driver loop of src/RP_FVM.c:179-191, "
        "dt + step timed as src/RP_FVM.c:181-216, IO excluded) on the GPU host's cores, then the "
        "FP64 form on the GPU through the same loop (one dt read back per step) and through the "
        "device-side loop (rpfvm_advance_to: dt, clipping and clock on the device).", "",
        "| run | steps | seconds | Mcell-updates/s | cores |", "|---|---:|---:|---:|---:|",
        f"| reference CPU/OpenMP step | {d['steps']} | {c['seconds']:.2f} | {c['value']:.3f} | "
        f"{c['cores']} ({h['cpu_model']}, {h['cpus_visible']} CPUs visible, OMP_NUM_THREADS="
        f"{h['omp_num_threads_env']}) |",
        f"| FP64 form, one MI355X, host-driven loop | {g['steps']} | {g['seconds_wall']:.2f} | "
        f"{g['value']:.1f} | |"] + ([
        f"| FP64 form, one MI355X, device-side loop | {dl['steps']} | {dl['seconds_wall']:.2f} | "
        f"{dl['value']:.1f} | |"] if (dl := d.get("gpu_fp64_device_loop")) else []) + ["",
        f"Bit-identical to the reference: dts {b['dts']}, final state {b['state']}, final time "
        f"{b['t']}" + (f"; device-side loop: final state {dl['state_bit_identical']}, final "
                       f"time {dl['t']}." if dl else ".") + "  At 200^2 the step's arithmetic "
        "is a small part of a step: the host-driven loop pays one launch sequence and one host "
        "round trip per step, the device-side loop the launch sequence only."]) + "\n"


def ll3_markdown(d: dict, command: str) -> str:
    """The Lax-Liu 3 full-length accuracy record from `bench.py --ll3-full`'s JSON line."""
    n = d["workload"].rsplit("-", 1)[1]
    f = d["fp64"]
    lines = [f"# Lax-Liu configuration 3, {n}, to t = {d['t_end']}: accuracy of the mixed form",
             "", "Command (from the repository root):", "", "```", command, "```", "",
             "Four-quadrant Riemann problem (interface x = y = 0.8), WENO5-char + HLLC, SSP-RK3, "
             "outflow, one MI355X.  FP64: "
             f"{f['steps']} steps to t = {f['t']:.6f}, {f['wall_s']:.2f} s, faults {f['faults']}.",
             "", "| mixed run | steps | wall s | L1(rho) | rel | Linf(rho) | L1(p) | rel | Linf(p) |",
             "|---|---:|---:|---:|---:|---:|---:|---:|---:|"]
    for label, k in (("own dts", "mixed"), ("FP64 dts", "mixed_replay")):
        r = d[k]
        lines.append(f"| {label} | {r['steps']} | {r['wall_s']:.2f} | {r['l1_rho']:.3e} | "
                     f"{r['l1_rho_rel']:.3e} | {r['linf_rho']:.3e} | {r['l1_p']:.3e} | "
                     f"{r['l1_p_rel']:.3e} | {r['linf_p']:.3e} |")
    return "\n".join(lines) + "\n"


def solvers_markdown(d: dict, command: str) -> str:
    """The per-solver Noh wall-heating record from `bench.py --noh-solvers`'s JSON line."""
    n = d["workload"].split("-")[1]
    lines = [f"# Noh {n} to t = {d['t_end']}: wall heating per Riemann solver and form", "",
             "Command (from the repository root):", "", "```", command, "```", "",
             "Noh implosion, gamma 5/3, p0 = 1e-6, WENO5-char, SSP-RK3, CFL 0.4, positivity "
             "safeguard on, one MI355X.  Exact solution: rho = 16 behind the shock (r < t/3); "
             "the wall-heating deficit is 1 - rho/16 at the origin (corner cell) and over "
             "r < 0.05.  Fault bits (RPFVM_FAULT_*): 1 = an HLL face or RoeHC face took the "
             "reference's bad-state skip, 2 = non-finite dt, 4 = the mixed form recomputed "
             "faces in FP64 (informational), 8 = the safeguard left a cell with rho <= 0 or "
             "p <= 0.  Each form runs on its own dts; L1 is the mixed run's density distance "
             "from the FP64 run of the same solver.", "",
             "| solver | scheme | form | steps | faults | rho(corner cell) | deficit corner | "
             "deficit r<0.05 | rho plateau 0.1<r<0.15 | L1(rho) vs FP64 | Mcell-updates/s |",
             "|---|---:|---|---:|---:|---:|---:|---:|---:|---:|---:|"]
    for r in d["solvers"]:
        if r["finite"]:
            m = (f"{r['rho_corner']:.4f} | {r['deficit_corner']:+.4f} | "
                 f"{r['deficit_r_lt_0.05']:+.5f} | {r['rho_plateau_0.1_0.15']:.4f}")
        else:
            m = "non-finite | | |"
        l1 = f"{r['l1_rho_vs_fp64']:.2e}" if "l1_rho_vs_fp64" in r else ""
        lines.append(f"| {r['solver']} | {r['scheme']} | {r.get('form', 'fp64')} | {r['steps']} "
                     f"| {r['faults']} | {m} | {l1} | {r['mcell_updates_per_s']:.0f} |")
    return "\n".join(lines) + "\n"


def _merge(name: str, key: str, value) -> None:
    try:
        with open(TRAFFIC_JSON) as fh:
            allt = json.load(fh)
    except (OSError, ValueError):
        allt = {}
    allt.setdefault(name, {})[key] = value
    with open(TRAFFIC_JSON, "w") as fh:
        json.dump(allt, fh, indent=1, sort_keys=True)


def load_traffic(workload: str) -> dict:
    """The committed per-dispatch traffic figures for `workload` ({} if none)."""
    try:
        with open(TRAFFIC_JSON) as fh:
            return json.load(fh).get(workload, {})
    except (OSError, ValueError):
        return {}


def main(argv):
    if len(argv) >= 4 and argv[0] == "trace":
        _, d, name, command = argv[:4]
        path = os.path.join(PROFILES, f"{name}_kernel_trace.md")
        with open(path, "w") as fh:
            fh.write(trace_markdown(d, name, command))
        print(path)
        return 0
    if len(argv) >= 8 and argv[0] == "traffic":
        _, fd, wd, cfd, cwd, cal, name, commands = argv[:8]
        t = traffic(fd, wd, cfd, cwd, int(cal))
        path = os.path.join(PROFILES, f"{name}_hbm_traffic.md")
        with open(path, "w") as fh:
            fh.write(traffic_markdown(t, name, commands))
        _merge(name, "calibration", t["calibration"])
        _merge(name, "kernels", t["kernels"])
        print(path)
        return 0
    if len(argv) >= 4 and argv[0] == "accuracy":
        _, jpath, name, command = argv[:4]
        with open(jpath) as fh:
            d = json.loads(fh.read().strip().splitlines()[-1])
        path = os.path.join(PROFILES, f"{name}_accuracy.md")
        with open(path, "w") as fh:
            fh.write(accuracy_markdown(d, command))
        _merge(d["workload"].replace("-", "").split("x")[0], "accuracy_full", accuracy_summary(d))
        print(path)
        return 0
    if len(argv) >= 4 and argv[0] in ("dmr", "c1"):
        kind, jpath, name, command = argv[:4]
        with open(jpath) as fh:
            d = json.loads(fh.read().strip().splitlines()[-1])
        path = os.path.join(PROFILES, f"{name}_{'accuracy' if kind == 'dmr' else 'record'}.md")
        with open(path, "w") as fh:
            fh.write(dmr_markdown(d, command) if kind == "dmr" else c1_markdown(d, command))
        print(path)
        return 0
    if len(argv) >= 4 and argv[0] == "ll3":
        _, jpath, name, command = argv[:4]
        with open(jpath) as fh:
            d = json.loads(fh.read().strip().splitlines()[-1])
        path = os.path.join(PROFILES, f"{name}_accuracy.md")
        with open(path, "w") as fh:
            fh.write(ll3_markdown(d, command))
        print(path)
        return 0
    if len(argv) >= 4 and argv[0] == "solvers":
        _, jpath, name, command = argv[:4]
        with open(jpath) as fh:
            d = json.loads(fh.read().strip().splitlines()[-1])
        path = os.path.join(PROFILES, f"{name}_solvers.md")
        with open(path, "w") as fh:
            fh.write(solvers_markdown(d, command))
        print(path)
        return 0
    if len(argv) >= 4 and argv[0] == "valu":
        _, d, name, command = argv[:4]
        v = valu(d)
        path = os.path.join(PROFILES, f"{name}_valu_counters.md")
        with open(path, "w") as fh:
            fh.write(valu_markdown(v, name, command))
        _merge(name, "valu", v)
        print(path)
        return 0
    print(__doc__, file=sys.stderr)
    return 2


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
