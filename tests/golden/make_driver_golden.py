"""Golden output files of the reference program for the C driver's tests.

    make -C oracle ref && python tests/golden/make_driver_golden.py

Runs oracle/_ref/rp_fvm (the reference's src/RP_FVM.c, compiled from its own sources)
on small control files, one OpenMP thread (its 2-D Noh initialiser draws rand() inside
a parallel loop), and stores every output file (.plt / .dat) it writes, by name, in
tests/golden/driver.npz together with the control-file text (output_dir left out); for
the Precision_Test runs, which write no files, the printed errors and error table.
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
PROG = os.path.join(ROOT, "oracle", "_ref", "rp_fvm")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "driver.npz")

# name -> control-file keys (a small grid of each family; end times are the cases' own)
CASES = {
    "riemann3_hllc": dict(TestCase="2D_Riemann_Case3", L_nx=24, L_ny=20, scheme=2, Control_output=2),
    "riemann1_roe": dict(TestCase="2D_Riemann_Case1", L_nx=16, L_ny=18, scheme=3, Control_output=1),
    "riemann5_hll_rk2": dict(TestCase="2D_Riemann_Case5", L_nx=20, L_ny=16, scheme=1, Time_ADM=2,
                             Control_output=2),
    "sod_1d_llf": dict(TestCase="1D_Sod_Shocktube", L_nx=64, L_ny=1, scheme=10, Control_output=2),
    "dmr_hllc": dict(TestCase="Double_Mach_Reflection", L_nx=48, L_ny=12, scheme=2, Control_output=1),
    "blast_hllc": dict(TestCase="Blast_Wave", L_nx=24, L_ny=24, scheme=2, Control_output=2),
    "rt_gravity": dict(TestCase="Rayleigh_Taylor", L_nx=8, L_ny=32, scheme=2, Source="true",
                       Gravity=1.0, M_gamma=1.4, Control_output=1),
    "kh_hllc": dict(TestCase="Kelvin_Helmholtz", L_nx=16, L_ny=16, scheme=2, Control_output=1, CFL=0.5),
    "noh2d_hllc": dict(TestCase="2D_Noh_Problem", L_nx=16, L_ny=16, scheme=2, M_gamma=5.0 / 3.0,
                       Control_output=1),
    "oedc_roe": dict(TestCase="Odd_Even_Decoupling", L_nx=40, L_ny=8, scheme=3, Control_output=1),
    "backstep_rk1_steps": dict(TestCase="Backward_Step", L_nx=24, L_ny=8, scheme=2, Time_ADM=1,
                               Control_Compution=1, Control_output=100),
}
# Precision_Test runs (src/RP_FVM.c:974-1072): no output files; the golden is the
# program's stdout from the first error line on (errors, orders, the error table)
PRECISION_CASES = {
    "precision_hllc": dict(TestCase="Precision_Test", L_nx=10, L_ny=1, scheme=2),
    "precision_hllhc": dict(TestCase="Precision_Test", L_nx=10, L_ny=1, scheme=11),
    "precision_roehc_rk2": dict(TestCase="Precision_Test", L_nx=10, L_ny=1, scheme=33, Time_ADM=2),
}
STDOUT_FROM = "L1 Error = "
DEFAULTS = dict(Time_ADM=3, M_gamma=1.4, Source="false", Gravity=1.0, CFL=0.4, Tmax=1.0,
                Control_Compution=0)


def control_text(keys, output_dir):
    d = dict(DEFAULTS)
    d.update(keys)
    lines = ["# generated control file"] + [f"{k}: {v}" for k, v in d.items()]
    lines.append(f"output_dir: {output_dir}")
    return "\n".join(lines) + "\n"


def main():
    if not os.path.exists(PROG):
        sys.exit(f"{PROG} missing: run `make -C oracle ref` first")
    data = {}
    for name, keys in CASES.items():
        with tempfile.TemporaryDirectory() as td:
            out = os.path.join(td, "out")
            cf = os.path.join(td, "ControlFile.txt")
            with open(cf, "w") as f:
                f.write(control_text(keys, out))
            env = dict(os.environ, OMP_NUM_THREADS="1")
            subprocess.run([PROG, cf], cwd=td, env=env, check=True, stdout=subprocess.DEVNULL,
                           timeout=600)
            files = sorted(os.listdir(out))
            data[name + "__control"] = np.array(control_text(keys, "@OUT@"))
            data[name + "__files"] = np.array(files)
            for k, fn in enumerate(files):
                with open(os.path.join(out, fn)) as f:
                    data[f"{name}__{k}"] = np.array(f.read())
        print(name, files, flush=True)
    for name, keys in PRECISION_CASES.items():
        with tempfile.TemporaryDirectory() as td:
            cf = os.path.join(td, "ControlFile.txt")
            with open(cf, "w") as f:
                f.write(control_text(keys, os.path.join(td, "out")))
            env = dict(os.environ, OMP_NUM_THREADS="1")
            r = subprocess.run([PROG, cf], cwd=td, env=env, check=True, capture_output=True,
                               text=True, timeout=600)
            out = r.stdout[r.stdout.index(STDOUT_FROM):]
            data[name + "__control"] = np.array(control_text(keys, "@OUT@"))
            data[name + "__stdout"] = np.array(out)
        print(name, out.splitlines()[-3], flush=True)
    np.savez_compressed(OUT, **data)
    print(f"{OUT}: {os.path.getsize(OUT) / 1024:.1f} KiB")


if __name__ == "__main__":
    main()
