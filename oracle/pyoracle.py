"""ctypes binding of the CPU oracle (oracle/liboracle_rpfvm.so).  Test infrastructure:
imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg."""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_rpfvm.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")

_lib = None


def build(reference_dir: str | None = None) -> None:
    """Compile the oracle (and, when the reference sources are present, the
    reference harness) with oracle/Makefile."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    ref = reference_dir or "/root/reference"
    if os.path.isdir(os.path.join(ref, "include")):
        subprocess.run(["make", "-s", "-C", HERE, "ref", f"REF={ref}"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library not built: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        vp = ctypes.c_void_p
        L.orc_delta_t.restype = ctypes.c_double
        L.orc_delta_t.argtypes = [vp, dp]
        L.orc_boundary_conditions.argtypes = [vp, dp]
        L.orc_boundary_conditions_t.argtypes = [vp, dp, ctypes.c_double]
        L.orc_space_operator.argtypes = [vp, dp, dp, dp, dp]
        L.orc_time_advance.argtypes = [vp, dp, ctypes.c_double]
        L.orc_time_advance_t.argtypes = [vp, dp, ctypes.c_double, ctypes.c_double]
        L.orc_advance_to.argtypes = [vp, dp, ctypes.c_double, ctypes.c_int, dp, dp]
        L.orc_reconstruct.argtypes = [vp, dp, ctypes.c_int, dp, dp]
        L.orc_riemann.argtypes = [vp, ctypes.c_int, dp, dp, dp]
        L.orc_weno5_points.argtypes = [ctypes.c_int, dp, dp]
        L.orc_eigen.argtypes = [vp, dp, ctypes.c_int, dp, dp]
        L.orc_threads.restype = ctypes.c_int
        L.orc_last_repairs.restype = ctypes.c_long
        _lib = L
    return _lib


def _ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _cfgp(cfg):
    return ctypes.cast(ctypes.pointer(cfg), ctypes.c_void_p)


def _check(rc):
    if rc < 0:
        raise RuntimeError(f"oracle error {rc}")
    return rc


def delta_t(cfg, U):
    U = np.ascontiguousarray(U, dtype=np.float64)
    return lib().orc_delta_t(_cfgp(cfg), _ptr(U))


def boundary_conditions(cfg, U, t=0.0):
    """In place at clock t (read by the time-dependent sides); returns U."""
    _check(lib().orc_boundary_conditions_t(_cfgp(cfg), _ptr(U), ctypes.c_double(t)))
    return U


def space_operator(cfg, U):
    """Ghost fill U in place; returns (L, F, G, faults)."""
    L = np.zeros_like(U)
    F = np.zeros_like(U)
    G = np.zeros_like(U)
    rc = _check(lib().orc_space_operator(_cfgp(cfg), _ptr(U), _ptr(L), _ptr(F), _ptr(G)))
    return L, F, G, rc


def time_advance(cfg, U, dt, t=0.0):
    """One RK step in place from clock t; returns fault bits."""
    return _check(lib().orc_time_advance_t(_cfgp(cfg), _ptr(U), ctypes.c_double(dt),
                                           ctypes.c_double(t)))


def advance_to(cfg, U, t_stop, max_steps, t=0.0):
    tt = np.array([t], dtype=np.float64)
    dts = np.zeros(max(1, max_steps), dtype=np.float64)
    n = _check(lib().orc_advance_to(_cfgp(cfg), _ptr(U), ctypes.c_double(t_stop),
                                    int(max_steps), _ptr(tt), _ptr(dts)))
    return float(tt[0]), dts[:n]


def run_steps(cfg, U, steps):
    """`steps` steps with dt from Get_Delta_T_2D each step (no clipping); returns dts."""
    dts = []
    for _ in range(steps):
        dt = delta_t(cfg, U)
        dts.append(dt)
        time_advance(cfg, U, dt)
    return np.array(dts)


def reconstruct(cfg, U, direction):
    CL = np.zeros_like(U)
    CR = np.zeros_like(U)
    _check(lib().orc_reconstruct(_cfgp(cfg), _ptr(U), int(direction), _ptr(CL), _ptr(CR)))
    return CL, CR


def riemann(cfg, direction, Ls, Rs, F=None):
    F = np.zeros_like(Ls) if F is None else F
    rc = _check(lib().orc_riemann(_cfgp(cfg), int(direction), _ptr(Ls), _ptr(Rs), _ptr(F)))
    return F, rc


def weno5_points(stencils):
    s = np.ascontiguousarray(stencils, dtype=np.float64)
    out = np.zeros((s.shape[0], 4), dtype=np.float64)
    lib().orc_weno5_points(int(s.shape[0]), _ptr(s), _ptr(out))
    return out


def eigen(cfg, U, direction):
    R, C = cfg.R, cfg.C
    EL = np.zeros((4, 4, R, C))
    ER = np.zeros((4, 4, R, C))
    _check(lib().orc_eigen(_cfgp(cfg), _ptr(np.ascontiguousarray(U)), int(direction),
                           _ptr(EL), _ptr(ER)))
    return EL, ER


def threads():
    return int(lib().orc_threads())


def last_repairs():
    """Cells the positivity safeguard repaired during the last time_advance."""
    return int(lib().orc_last_repairs())


# ------------------------------------------------------------ reference harness
def harness_job(mode, cfg=None, **extra):
    """Text of a job file for oracle/_ref/ref_harness (see oracle/ref_harness.c)."""
    lines = [f"mode {mode}"]
    if cfg is not None:
        d = cfg.as_dict()
        lines += [f"nx {d['nx']}", f"ny {d['ny']}", f"time_order {d['time_order']}",
                  f"scheme {d['scheme']}", f"recon {d['recon_order']}",
                  f"characteriz {d['characteristic']}", f"source {d['source']}",
                  f"bc_left {d['bc'][0]}", f"bc_right {d['bc'][1]}",
                  f"bc_bottom {d['bc'][2]}", f"bc_top {d['bc'][3]}",
                  f"gamma {d['gamma']!r}", f"cfl {d['cfl']!r}", f"dx {d['dx']!r}",
                  f"dy {d['dy']!r}", f"gravity {d['gravity']!r}"]
        for k in range(4):
            lines.append(f"inflow{k} {d['inflow_state'][k]!r}")
        for s in range(4):
            for k in range(4):
                lines.append(f"fixed{s}{k} {d['fixed_state'][s][k]!r}")
    for k, v in extra.items():
        lines.append(f"{k} {v!r}" if isinstance(v, float) else f"{k} {v}")
    return "\n".join(lines) + "\n"


def harness_run(cfg, U, t_end, *, max_steps=0, threads=None):
    """The reference's own time loop (src/RP_FVM.c:179-191, next output time = Tmax =
    t_end) through oracle/_ref/ref_harness mode "run".  Returns (dts, t, U, record): the
    dt of every step, the final time, the final state (whole array, ghosts included) and
    the harness's timing record (seconds of dt + step, IO excluded; thread count)."""
    import json
    import tempfile
    if not os.path.exists(REF_HARNESS):
        raise RuntimeError(f"reference harness not built: {REF_HARNESS} (make -C oracle ref)")
    env = dict(os.environ)
    if threads:
        env.update(OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="true")
    U = np.ascontiguousarray(U, dtype="<f8")
    with tempfile.TemporaryDirectory() as td:
        job = os.path.join(td, "job.txt")
        with open(job, "w") as f:
            f.write(harness_job("run", cfg, t_end=float(t_end), steps=int(max_steps)))
        U.tofile(os.path.join(td, "in.bin"))
        r = subprocess.run([REF_HARNESS, job, os.path.join(td, "in.bin"),
                            os.path.join(td, "out.bin")], stdout=subprocess.PIPE, text=True,
                           env=env, check=True)  # its progress lines go to our stderr
        raw = np.fromfile(os.path.join(td, "out.bin"), dtype="<f8")
    n, t = int(raw[0]), float(raw[1])
    rec = json.loads(r.stdout.strip().splitlines()[-1])
    return raw[2:2 + n].copy(), t, raw[2 + n:].reshape(U.shape).copy(), rec
