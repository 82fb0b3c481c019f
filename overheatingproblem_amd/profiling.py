"""Summaries of rocprofv3 CSV output for the committed profile records.

    python -m overheatingproblem_amd.profiling trace   <trace_dir> <name> "<command>"
    python -m overheatingproblem_amd.profiling traffic <fetch_dir> <write_dir> \
        <cal_fetch_dir> <cal_write_dir> <cal_bytes> <name> "<commands>"
    python -m overheatingproblem_amd.profiling valu    <pmc_dir> <name> "<command>"
    python -m overheatingproblem_amd.profiling accuracy <noh_full.json> <name> "<command>"
    python -m overheatingproblem_amd.profiling solvers <noh_solvers.json> <name> "<command>"
    python -m overheatingproblem_amd.profiling ll3     <ll3_full.json> <name> "<command>"

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


def valu(pmc_dir: str) -> dict:
    """Per-dispatch SQ instruction counts (SQ_INSTS_VALU etc.) of each kernel."""
    out = {}
    for k, cs in counter_means(pmc_dir).items():
        if k.startswith("__amd") or "SQ_INSTS_VALU" not in cs:
            continue
        rec = {c.lower() + "_per_dispatch": v for c, (v, _) in cs.items()}
        rec["valu_wave_instr_per_dispatch"] = cs["SQ_INSTS_VALU"][0]
        rec["dispatches"] = cs["SQ_INSTS_VALU"][1]
        out[k] = rec
    return out


def valu_markdown(v: dict, name: str, command: str) -> str:
    cols = sorted({c for rec in v.values() for c in rec if c.endswith("_per_dispatch")})
    lines = [f"# SQ instruction counters per dispatch: {name}", "",
             "Command (from the repository root):", "", "```", command, "```", "",
             "| kernel | dispatches | " + " | ".join(c.replace("_per_dispatch", "") for c in cols)
             + " |", "|---|---:|" + "---:|" * len(cols)]
    for k, rec in v.items():
        lines.append(f"| {k} | {rec['dispatches']} | "
                     + " | ".join(f"{rec.get(c, 0):.4g}" for c in cols) + " |")
    return "\n".join(lines) + "\n"


def accuracy_markdown(d: dict, command: str) -> str:
    """The Noh full-length accuracy record from `bench.py --noh-full`'s JSON line."""
    runs = (("FP64", "fp64"), ("mixed (own dts)", "mixed"), ("mixed (FP64 dts)", "mixed_replay"),
            ("FP64, rho perturbed 1e-12", "fp64_perturbed"))
    n = d["workload"].split("-")[1]
    lines = [f"# Noh {n} to t = {d['t_end']}: accuracy of the mixed form", "",
             "Command (from the repository root):", "", "```", command, "```", "",
             "Noh implosion, gamma 5/3, p0 = 1e-6, WENO5-char + HLLC, SSP-RK3, CFL 0.4, "
             "positivity safeguard on, one MI355X.  Exact solution: rho = 16 behind the shock "
             "(r < t/3); the wall-heating deficit is 1 - rho/16 near the origin.", "",
             "| run | steps | wall s | rho(corner cell) | deficit corner | deficit r<0.05 | "
             "rho plateau 0.1<r<0.15 |", "|---|---:|---:|---:|---:|---:|---:|"]
    for label, k in runs:
        r = d[k]
        lines.append(f"| {label} | {r['steps']} | {r['wall_s']:.2f} | {r['rho_corner']:.4f} | "
                     f"{r['deficit_corner']:+.4f} | {r['deficit_r_lt_0.05']:+.5f} | "
                     f"{r['rho_plateau_0.1_0.15']:.4f} |")
    lines += ["", "Differences from the FP64 run (interior, absolute and relative to mean |FP64|):",
              "", "| compared with FP64 | L1(rho) | rel | Linf(rho) | L1(p) | rel | Linf(p) |",
              "|---|---:|---:|---:|---:|---:|---:|"]
    for label, k in runs[1:]:
        c = d[k + "_vs_fp64"]
        lines.append(f"| {label} | {c['l1_rho']:.3e} | {c['l1_rho_rel']:.3e} | {c['linf_rho']:.3f} "
                     f"| {c['l1_p']:.3e} | {c['l1_p_rel']:.3e} | {c['linf_p']:.3f} |")
    ref = d["fp64_perturbed_vs_fp64"]["l1_rho"]
    ratios = ", ".join(f"{label} {d[k + '_vs_fp64']['l1_rho'] / ref:.2f}" for label, k in runs[1:3])
    lines += ["", "Reading: the Noh solution amplifies round-off at the origin (wall heating) and at "
              "the shock, so over the full run FP64 itself, started from a density perturbed by "
              "1e-12, ends this far from the unperturbed FP64 run; the mixed form's L1(rho) "
              f"distance from FP64, in units of that FP64 self-distance: {ratios}."]
    return "\n".join(lines) + "\n"


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
    lines = [f"# Noh {n} to t = {d['t_end']}: wall heating per Riemann solver (FP64 form)", "",
             "Command (from the repository root):", "", "```", command, "```", "",
             "Noh implosion, gamma 5/3, p0 = 1e-6, WENO5-char, SSP-RK3, CFL 0.4, positivity "
             "safeguard on, one MI355X.  Exact solution: rho = 16 behind the shock (r < t/3); "
             "the wall-heating deficit is 1 - rho/16 at the origin (corner cell) and over "
             "r < 0.05.  Fault bits (RPFVM_FAULT_*): 1 = an HLL face or RoeHC face took the "
             "reference's bad-state skip, 2 = non-finite dt, 8 = the safeguard left a cell with "
             "rho <= 0 or p <= 0.", "",
             "| solver | scheme | steps | t | faults | rho(corner cell) | deficit corner | "
             "deficit r<0.05 | rho plateau 0.1<r<0.15 | Mcell-updates/s |",
             "|---|---:|---:|---:|---:|---:|---:|---:|---:|---:|"]
    for r in d["solvers"]:
        if r["finite"]:
            m = (f"{r['rho_corner']:.4f} | {r['deficit_corner']:+.4f} | "
                 f"{r['deficit_r_lt_0.05']:+.5f} | {r['rho_plateau_0.1_0.15']:.4f}")
        else:
            m = "non-finite | | |"
        lines.append(f"| {r['solver']} | {r['scheme']} | {r['steps']} | {r['t']:.4f} | {r['faults']} "
                     f"| {m} | {r['mcell_updates_per_s']:.0f} |")
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
