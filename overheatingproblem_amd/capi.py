"""ctypes binding of the gfx950 library (overheatingproblem_amd/lib/librpfvm.so).

``Solver`` mirrors the reference's step interface (the functions its driver calls,
src/RP_FVM.c:179-209):

    Get_Delta_T_2D           (include/function.h:144)        -> Solver.get_delta_t()
    RK1/RK2/RK3_TimeAd       (include/time_advance.h:643-882) -> Solver.time_advance(dt)
    Boundary_Conditions      (include/Boundary_Condition.h:963) -> Solver.boundary_conditions()
    Flux_Reconstruction_RP + Space_Discrete_Item
                             (include/CFD_convection.h:38, include/scheme.h:3758)
                                                               -> Solver.space_operator()

There is no CPU fallback: without the built library (or without a GPU) every call
raises.
"""
from __future__ import annotations

import atexit
import ctypes
import os
import weakref

import numpy as np

from .config import Config

GHOST = 4  # ghost layers per side

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "librpfvm.so")

# Every entry point include/rpfvm.h declares.
EXPORTS = (
    "rpfvm_config_default", "rpfvm_create", "rpfvm_destroy", "rpfvm_set_state",
    "rpfvm_get_state", "rpfvm_get_delta_t", "rpfvm_time_advance", "rpfvm_boundary_conditions",
    "rpfvm_space_operator", "rpfvm_riemann", "rpfvm_advance_to", "rpfvm_sync", "rpfvm_get_faults",
    "rpfvm_device_state", "rpfvm_stream", "rpfvm_set_kernel_timing", "rpfvm_kernel_times",
    "rpfvm_last_error", "rpfvm_version", "rpfvm_set_time", "rpfvm_get_time",
    "rpfvm_batch_create", "rpfvm_batch_destroy", "rpfvm_batch_set_state", "rpfvm_batch_get_state",
    "rpfvm_batch_step", "rpfvm_batch_advance_to", "rpfvm_batch_status", "rpfvm_batch_sync",
    "rpfvm_batch_tune", "rpfvm_batch_set_tuning", "rpfvm_batch_get_tuning", "rpfvm_batch_stream",
    "rpfvm_delta_t_noalloc",
)

STAGING_DIRECT, STAGING_LDS, STAGING_SWEEPS = 0, 1, 2


class BatchTuning(ctypes.Structure):
    """rpfvm_batch_tuning (include/rpfvm.h)."""
    _fields_ = [("staging", ctypes.c_int), ("rows", ctypes.c_int), ("waves_per_cu", ctypes.c_int),
                ("use_graph", ctypes.c_int), ("ms_per_stage", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}

KERNEL_CLASSES = ("bc", "flux_x", "flux_y", "update", "dt")


class RpfvmError(RuntimeError):
    pass


_lib = None

# Every open Solver / Batch.  Destroyed at interpreter exit, before the HIP runtime's own
# teardown (C-level exit handlers run after Python's), so that no stream, event, graph or
# device allocation of a context outlives the runtime that owns it: a context left to
# __del__ during interpreter shutdown may find the module globals already cleared, skip
# rpfvm_destroy, and leave its HIP objects to the runtime's exit path.
_live = weakref.WeakSet()


def _close_all() -> None:
    for obj in list(_live):
        try:
            obj.close()
        except Exception:
            pass


atexit.register(_close_all)


def lib():
    """Load librpfvm.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RpfvmError(f"{LIB_PATH} is missing: build it with python -m overheatingproblem_amd.build")
    L = ctypes.CDLL(LIB_PATH)
    vp, dp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    szp = ctypes.POINTER(ctypes.c_size_t)
    L.rpfvm_config_default.argtypes = [ctypes.POINTER(Config)]
    L.rpfvm_config_default.restype = None
    L.rpfvm_create.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(vp)]
    L.rpfvm_destroy.argtypes = [vp]
    L.rpfvm_set_state.argtypes = [vp, dp]
    L.rpfvm_get_state.argtypes = [vp, dp]
    L.rpfvm_get_delta_t.argtypes = [vp, dp]
    L.rpfvm_time_advance.argtypes = [vp, ctypes.c_double]
    L.rpfvm_boundary_conditions.argtypes = [vp]
    L.rpfvm_space_operator.argtypes = [vp, dp, dp, dp]
    L.rpfvm_riemann.argtypes = [vp, ctypes.c_int, dp, dp, dp]
    L.rpfvm_advance_to.argtypes = [vp, ctypes.c_double, ctypes.c_int, dp, ip]
    L.rpfvm_sync.argtypes = [vp]
    L.rpfvm_get_faults.argtypes = [vp, ip]
    L.rpfvm_device_state.argtypes = [vp, ctypes.POINTER(ctypes.c_void_p), szp, szp]
    L.rpfvm_stream.argtypes = [vp, ctypes.POINTER(vp)]
    L.rpfvm_set_kernel_timing.argtypes = [vp, ctypes.c_int]
    L.rpfvm_kernel_times.argtypes = [vp, dp, ip]
    L.rpfvm_set_time.argtypes = [vp, ctypes.c_double]
    L.rpfvm_get_time.argtypes = [vp, dp]
    tp = ctypes.POINTER(BatchTuning)
    L.rpfvm_batch_create.argtypes = [ctypes.c_int, ctypes.POINTER(Config), ctypes.POINTER(vp)]
    L.rpfvm_batch_destroy.argtypes = [vp]
    L.rpfvm_batch_set_state.argtypes = [vp, ctypes.c_int, dp]
    L.rpfvm_batch_get_state.argtypes = [vp, ctypes.c_int, dp]
    L.rpfvm_batch_step.argtypes = [vp, ctypes.c_double]
    L.rpfvm_batch_advance_to.argtypes = [vp, ctypes.c_double, ctypes.c_int, ip]
    L.rpfvm_batch_status.argtypes = [vp, dp, ip, ip]
    L.rpfvm_batch_sync.argtypes = [vp]
    L.rpfvm_batch_tune.argtypes = [vp, tp]
    L.rpfvm_batch_set_tuning.argtypes = [vp, tp]
    L.rpfvm_batch_get_tuning.argtypes = [vp, tp]
    L.rpfvm_batch_stream.argtypes = [vp, ctypes.POINTER(vp)]
    L.rpfvm_delta_t_noalloc.argtypes = [ctypes.POINTER(Config), vp, ctypes.c_size_t, ctypes.c_size_t,
                                        vp, vp, vp, vp]
    L.rpfvm_last_error.restype = ctypes.c_char_p
    L.rpfvm_version.restype = ctypes.c_char_p
    for name in EXPORTS:
        if name not in ("rpfvm_config_default", "rpfvm_last_error", "rpfvm_version"):
            getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def _check(rc: int) -> int:
    if rc < 0:
        msg = lib().rpfvm_last_error().decode(errors="replace")
        raise RpfvmError(f"rpfvm error {rc}: {msg}")
    return rc


def _dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Solver:
    """One device context of the step (``rpfvm_ctx``)."""

    def __init__(self, cfg: Config):
        self.cfg = cfg.copy()
        self._ctx = ctypes.c_void_p()
        _check(lib().rpfvm_create(ctypes.byref(self.cfg), ctypes.byref(self._ctx)))
        _live.add(self)

    # context management ------------------------------------------------------
    def close(self):
        if self._ctx:
            ctx, self._ctx = self._ctx, ctypes.c_void_p()
            lib().rpfvm_destroy(ctx)
        _live.discard(self)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def shape(self):
        return self.cfg.shape

    # state -------------------------------------------------------------------
    def set_state(self, U: np.ndarray) -> None:
        U = np.ascontiguousarray(U, dtype=np.float64)
        if U.shape != self.shape:
            raise ValueError(f"state shape {U.shape} != {self.shape}")
        _check(lib().rpfvm_set_state(self._ctx, _dptr(U)))

    def get_state(self) -> np.ndarray:
        U = np.empty(self.shape, dtype=np.float64)
        _check(lib().rpfvm_get_state(self._ctx, _dptr(U)))
        return U

    # the reference's step interface -----------------------------------------
    def get_delta_t(self) -> float:
        dt = ctypes.c_double()
        _check(lib().rpfvm_get_delta_t(self._ctx, ctypes.byref(dt)))
        return dt.value

    def time_advance(self, dt: float) -> None:
        _check(lib().rpfvm_time_advance(self._ctx, ctypes.c_double(dt)))

    def boundary_conditions(self) -> None:
        _check(lib().rpfvm_boundary_conditions(self._ctx))

    def space_operator(self):
        """Returns (L, F, G) as host arrays in the reference layout."""
        L = np.zeros(self.shape)
        F = np.zeros(self.shape)
        G = np.zeros(self.shape)
        _check(lib().rpfvm_space_operator(self._ctx, _dptr(L), _dptr(F), _dptr(G)))
        return L, F, G

    def riemann(self, direction: int, Ls: np.ndarray, Rs: np.ndarray) -> np.ndarray:
        """The configured solver alone on host face states (rpfvm_riemann): the face
        fluxes of ``direction`` (1: x, 2: y) at the faces' left cells (entries off the
        face range are unspecified)."""
        Ls = np.ascontiguousarray(Ls, dtype=np.float64)
        Rs = np.ascontiguousarray(Rs, dtype=np.float64)
        if Ls.shape != self.shape or Rs.shape != self.shape:
            raise ValueError(f"face states must have shape {self.shape}")
        F = np.zeros(self.shape)
        _check(lib().rpfvm_riemann(self._ctx, int(direction), _dptr(Ls), _dptr(Rs), _dptr(F)))
        return F

    def set_time(self, t: float) -> None:
        """Clock read by the time-dependent boundary sides (rpfvm_set_time)."""
        _check(lib().rpfvm_set_time(self._ctx, ctypes.c_double(t)))

    def get_time(self) -> float:
        t = ctypes.c_double()
        _check(lib().rpfvm_get_time(self._ctx, ctypes.byref(t)))
        return t.value

    def advance_to(self, t_stop: float, max_steps: int, t: float = 0.0):
        tt = ctypes.c_double(t)
        n = ctypes.c_int(0)
        _check(lib().rpfvm_advance_to(self._ctx, ctypes.c_double(t_stop), int(max_steps),
                                      ctypes.byref(tt), ctypes.byref(n)))
        return tt.value, n.value

    def sync(self) -> int:
        """Wait for the stream; returns (and clears) the fault bits."""
        return _check(lib().rpfvm_sync(self._ctx))

    def device_state(self):
        base = ctypes.c_void_p()
        ld = ctypes.c_size_t()
        plane = ctypes.c_size_t()
        _check(lib().rpfvm_device_state(self._ctx, ctypes.byref(base), ctypes.byref(ld),
                                        ctypes.byref(plane)))
        return base.value, ld.value, plane.value

    def set_kernel_timing(self, enabled: bool) -> None:
        _check(lib().rpfvm_set_kernel_timing(self._ctx, int(bool(enabled))))

    def kernel_times(self) -> dict:
        """{class: (total_ms, launches)} since the last call; synchronises."""
        ms = (ctypes.c_double * len(KERNEL_CLASSES))()
        n = (ctypes.c_int * len(KERNEL_CLASSES))()
        _check(lib().rpfvm_kernel_times(self._ctx, ms, n))
        return {k: (ms[i], n[i]) for i, k in enumerate(KERNEL_CLASSES)}

    def stream(self) -> int:
        s = ctypes.c_void_p()
        _check(lib().rpfvm_stream(self._ctx, ctypes.byref(s)))
        return s.value or 0


class Batch:
    """Many independent problems advanced together (rpfvm_batch_*).  Each problem has
    its own Config; recon_order, characteristic, time_order and precision are shared.
    """

    def __init__(self, cfgs):
        L = lib()
        self.cfgs = [c.copy() for c in cfgs]
        arr = (Config * len(self.cfgs))(*self.cfgs)
        self._h = ctypes.c_void_p()
        _check(L.rpfvm_batch_create(len(self.cfgs), arr, ctypes.byref(self._h)))
        _live.add(self)

    def close(self):
        if getattr(self, "_h", None):
            h, self._h = self._h, None
            lib().rpfvm_batch_destroy(h)
        _live.discard(self)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.cfgs)

    def shape(self, k: int):
        c = self.cfgs[k]
        return (4, c.nx + 2 * GHOST, c.ny + 2 * GHOST)

    def set_state(self, k: int, U: np.ndarray) -> None:
        U = np.ascontiguousarray(U, dtype=np.float64)
        if U.shape != self.shape(k):
            raise ValueError(f"state of problem {k} must have shape {self.shape(k)}, got {U.shape}")
        _check(lib().rpfvm_batch_set_state(self._h, k, _dptr(U)))

    def get_state(self, k: int, out: np.ndarray | None = None) -> np.ndarray:
        """Problem k's state; the interior is current, ghost entries are those of `out`
        (zeros when not given)."""
        U = np.zeros(self.shape(k)) if out is None else out
        _check(lib().rpfvm_batch_get_state(self._h, k, _dptr(U)))
        return U

    def step(self, t_end: float) -> None:
        _check(lib().rpfvm_batch_step(self._h, float(t_end)))

    def advance_to(self, t_end: float, max_steps: int = 1 << 30) -> int:
        n = ctypes.c_int()
        _check(lib().rpfvm_batch_advance_to(self._h, float(t_end), int(max_steps), ctypes.byref(n)))
        return n.value

    def status(self):
        """(t, steps, faults) arrays, one entry per problem."""
        n = len(self.cfgs)
        t = np.zeros(n)
        st = np.zeros(n, dtype=np.int32)
        f = np.zeros(n, dtype=np.int32)
        _check(lib().rpfvm_batch_status(self._h, _dptr(t), st.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                        f.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return t, st, f

    def sync(self) -> int:
        return _check(lib().rpfvm_batch_sync(self._h))

    def tune(self) -> dict:
        t = BatchTuning()
        _check(lib().rpfvm_batch_tune(self._h, ctypes.byref(t)))
        return t.as_dict()

    def get_tuning(self) -> dict:
        t = BatchTuning()
        _check(lib().rpfvm_batch_get_tuning(self._h, ctypes.byref(t)))
        return t.as_dict()

    def set_tuning(self, **kw) -> None:
        t = BatchTuning()
        _check(lib().rpfvm_batch_get_tuning(self._h, ctypes.byref(t)))
        for k, v in kw.items():
            setattr(t, k, v)
        _check(lib().rpfvm_batch_set_tuning(self._h, ctypes.byref(t)))

    def stream(self) -> int:
        s = ctypes.c_void_p()
        _check(lib().rpfvm_batch_stream(self._h, ctypes.byref(s)))
        return s.value or 0


def version() -> str:
    return lib().rpfvm_version().decode()
